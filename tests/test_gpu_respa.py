"""Multiple time step (r-RESPA, no counterpart in the reference): nonbonded forces
every k inner steps with weight k, bonded terms every step.

Pinned by what the scheme fixes: with k = 1 it is the plain step loop bit for
bit; on an equilibrated united-atom melt in NVE its total energy (averaged over
whole cycles) drifts no more than a small multiple of the plain 1 fs loop's and
less than the plain loop run at the outer step (4 fs) over the same simulated
time; and each step evaluates the nonbonded force only when it should (the
nonbonded energy is constant inside a cycle).
"""
import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _ua(dt=0.001):
    s = system.ua_melt(nchains=20, nbeads=100, seed=5, avoid=0.6, dt=dt)
    e = Engine(s)
    workloads.equilibrate(e, s, stage_steps=100, final_steps=400)
    pos, vel = e.state()
    e.close()
    s.s, s.v = pos[:s.natoms].copy(), units.vel_from_reduced(vel)
    return s


def _energy_blocks(e, nblocks, block):
    """Mean total energy of consecutive blocks of steps (window sums of the step observables)."""
    out = []
    e.averages(reset=True)
    for _ in range(nblocks):
        e.step(block, graph=True)
        a = e.averages(reset=True)
        n = a[28]
        out.append((a[0:15].sum() + a[21] + a[22]) / n)
    return np.array(out)


def _drift(E, t):
    return np.polyfit(t, E, 1)[0]


def test_ratio_one_is_the_plain_loop():
    s = system.cg_melt(nchains=30, nbeads=50, seed=4, avoid=0.8, dt=0.001)
    a, b = Engine(s), Engine(s)
    b.set_respa(4)
    b.set_respa(1)
    a.step(35, graph=True)
    b.step(35, graph=True)
    sa, va = a.state()
    sb, vb = b.state()
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)


def test_respa_graph_equals_eager():
    """RESPA k = 4 (block period 40: lcm of the list interval 10 and k is 20, which holds an odd number of
    nonbonded evaluations, so it doubles) stepped through graphs from steps inside a cycle, in calls of 3, 47
    and 40 steps (blocks at phases 3 and 10): the same bits as eager steps."""
    s = _ua()                           # equilibrated: a raw random-walk melt blows up within 90 steps
    a, b = Engine(s), Engine(s)
    for e in (a, b):
        e.set_respa(4)
    a.step(90)
    for n in (3, 47, 40):
        b.step(n, graph=True)
    c = b.counters()
    print("counters", c, flush=True)
    assert c["period"] == 40 and c["long_block"] == 40
    assert c["graph_steps"] >= 80 and c["graph_launches"] >= 2
    sa, va = a.state()
    sb, vb = b.state()
    assert np.all(np.isfinite(sa)) and np.all(np.isfinite(va))
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)
    assert np.array_equal(a.obs()["vnb"], b.obs()["vnb"])


def test_nonbonded_only_on_cycle_steps():
    s = system.cg_melt(nchains=30, nbeads=50, seed=4, avoid=0.8, dt=0.001)
    e = Engine(s)
    e.set_respa(4)
    vnb = []
    for _ in range(12):
        e.step(1)
        vnb.append(float(e.obs()["vnb"].sum()))
    vnb = np.array(vnb)
    print("V_NB per step", vnb, flush=True)
    # steps 1, 5, 9 evaluate; 2-4, 6-8, 10-12 carry the last value
    for c in range(3):
        assert np.all(vnb[4 * c + 1:4 * c + 4] == vnb[4 * c])
        if c:
            assert vnb[4 * c] != vnb[4 * c - 1]


def test_energy_conservation_against_plain_steps():
    s = _ua()
    t_ps = 2.0
    res = {}
    for name, dt, k in (("plain_1fs", 0.001, 1), ("respa_4x1fs", 0.001, 4), ("plain_4fs", 0.004, 1)):
        e = Engine(s)
        e.set_thermostat(0, units.temp_to_reduced(s.control.temperature), 0.0, units.dt_to_reduced(dt))
        e.set_respa(k)
        nsteps = int(round(t_ps / dt))
        block = 40 if k > 1 or dt < 0.002 else 10
        E = _energy_blocks(e, nsteps // block, block)
        t = (np.arange(len(E)) + 0.5) * block * dt
        T = e.obs()["temp"] * units.TEMPSCALE
        e.close()
        assert np.all(np.isfinite(E))
        res[name] = (abs(_drift(E, t)), float(np.std(E - np.polyval(np.polyfit(t, E, 1), t))), T, float(E.mean()))
        print(name, "|drift| per ps", res[name][0], "scatter", res[name][1], "T", T, "E", res[name][3], flush=True)
    scale = abs(res["plain_1fs"][3])
    assert res["respa_4x1fs"][0] <= 5.0 * res["plain_1fs"][0] + 1e-4 * scale
    assert res["respa_4x1fs"][0] < res["plain_4fs"][0]
    assert abs(res["respa_4x1fs"][2] - res["plain_1fs"][2]) <= 0.1 * res["plain_1fs"][2]


def test_respa_matches_oracle(oracle_mod):
    """k = 4 on the equilibrated UA melt for three cycles, step by step against the FP64 oracle's RESPA (itself
    pinned by time reversibility and the impulse, test_respa_oracle.py): energies and virial within 1e-7 of
    their scale, final positions and velocities as in the plain short runs."""
    s = _ua()
    e = Engine(s)
    e.set_respa(4)
    o = oracle_mod.Oracle(s)
    o.set_respa(4)
    for k in range(12):
        e.step(1)
        o.run(1, k)
        og, oo = e.obs(), o.obs()
        eg = np.concatenate([og[q] for q in ("vnb", "vbond", "vangle", "vtors")])
        eo = np.concatenate([oo[q] for q in ("vnb", "vbond", "vangle", "vtors")])
        scale = np.abs(eo).sum()
        assert np.all(np.abs(eg - eo) <= 1e-7 * scale), (k + 1, eg, eo)
        assert np.all(np.abs(og["pt"] - oo["pt"]) <= 1e-7 * np.abs(oo["pt"][:3]).sum()), (k + 1, og["pt"], oo["pt"])
    sg, vg = e.state()
    so, vo, _ = o.state()
    ds = np.abs(sg - so).max()
    dv = float(np.linalg.norm(vg - vo) / np.linalg.norm(vo))
    print("RESPA k=4, 12 steps: max |ds|", ds, "velocity rel", dv, flush=True)
    assert ds <= 1e-6 and dv <= 1e-5


def test_respa_time_reversible():
    """On the device: 40 steps of k = 4 forward, then from (x_40, -v_{40+1/2}) 40 steps again return to x_0
    (FP32 forces, FP64 integration: to well under the 1e-3 nm a misplaced kick would leave)."""
    import copy
    s = _ua()
    e = Engine(s)
    e.set_respa(4)
    s0, _ = e.state()
    e.step(1)
    _, v_half = e.state()
    e.close()
    e = Engine(s)
    e.set_respa(4)
    e.step(40)
    sn, _ = e.state()
    e.step(1)
    _, vn1 = e.state()
    e.close()
    rev = copy.copy(s)
    rev.s, rev.v = sn[:s.natoms].copy(), units.vel_from_reduced(-vn1)
    r = Engine(rev)
    r.set_respa(4)
    r.step(40)
    sr, vr = r.state()
    ds = np.abs(sr - s0).max()
    dv = np.abs(vr + v_half).max() / np.abs(v_half).max()
    print("device RESPA k=4: 40 steps forward and back, max |x - x0|", ds, "velocity", dv, flush=True)
    assert ds <= 1e-6 and dv <= 1e-5
