"""The oracle's multiple time step (r-RESPA in leap-frog form; no counterpart in
the reference) pinned by what the scheme fixes, in FP64 on the CPU:

- time reversibility: N steps forward, then from (x_N, -v_{N+1/2}) N steps
  again return to x_0 (and -v_{1/2}) to round-off, when N is a whole number of
  cycles so the reversed run kicks at the same positions;
- the impulse: over one cycle of k steps the velocity change is k*dt*F_nb(x)/m
  at the first step plus dt*F_bonded at every step, so RESPA differs from the
  plain loop already after the first cycle, while k = 1 is the plain loop;
- steps inside a cycle report the last evaluation's nonbonded energy.
"""
import copy

import numpy as np

from ibisco_amd import system, units


def _melt():
    return system.cg_melt(nchains=12, nbeads=50, seed=41, avoid=0.8, dt=0.001)


def _reverse_run(oracle_mod, sy, k, n):
    o = oracle_mod.Oracle(sy)
    o.set_respa(k)
    s0, v0, _ = o.state()
    o.run(1, 0)
    _, v_half, _ = o.state()
    o = oracle_mod.Oracle(sy)
    o.set_respa(k)
    o.run(n, 0)
    sn, _, _ = o.state()
    o.run(1, n)                      # the next kick, from x_N: v_{N+1/2}
    _, vn1, _ = o.state()
    rev = copy.copy(sy)
    rev.s, rev.v = sn[:sy.natoms].copy(), units.vel_from_reduced(-vn1)
    r = oracle_mod.Oracle(rev)
    r.set_respa(k)
    r.run(n, 0)
    sr, vr, _ = r.state()
    return s0, v_half, sr, vr


def test_respa_oracle_time_reversible(oracle_mod):
    sy = _melt()
    for k, n in ((1, 24), (4, 24), (3, 24)):
        s0, v_half, sr, vr = _reverse_run(oracle_mod, sy, k, n)
        na = sy.natoms
        ds = np.abs(sr[:na] - s0[:na]).max()
        dv = np.abs(vr + v_half).max() / np.abs(v_half).max()
        print(f"k={k}: {n} steps forward and back, max |x - x0| {ds:.2e} nm, velocity {dv:.2e}")
        assert ds < 1e-9 and dv < 1e-8


def test_respa_impulse_and_reduction(oracle_mod):
    """k = 1 is the plain loop bit for bit; with k = 4 the first step's velocity change exceeds the plain step's
    by exactly 3 dt F_nb(x0)/m (F_nb from the full force minus the bonded-only force at the same positions),
    and steps 2-4 carry the first step's nonbonded energy."""
    sy = _melt()
    plain = oracle_mod.Oracle(sy)
    plain.run(4, 0)
    one = oracle_mod.Oracle(sy)
    one.set_respa(1)
    one.run(4, 0)
    assert np.array_equal(plain.state()[0], one.state()[0])
    full = oracle_mod.Oracle(sy)
    full.new_force()
    bonded_only = copy.deepcopy(sy)
    for t in bonded_only.nb_tables:
        t.v = np.zeros_like(np.asarray(t.v))
    b = oracle_mod.Oracle(bonded_only)
    b.new_force()
    f_nb = (full.forces() - b.forces())[:sy.natoms]
    p1 = oracle_mod.Oracle(sy)
    _, v0, _ = p1.state()
    p1.run(1, 0)
    four = oracle_mod.Oracle(sy)
    four.set_respa(4)
    four.run(1, 0)
    extra = four.state()[1] - p1.state()[1]
    dt = units.dt_to_reduced(sy.control.dt)
    m = units.mass_to_reduced(sy.type_mass[sy.itype])[:, None]
    expect = 3.0 * dt * f_nb / m
    err = np.linalg.norm(extra - expect) / np.linalg.norm(expect)
    print("RESPA k=4 first-step extra kick vs 3 dt F_nb/m: relative error", err)
    assert err < 1e-12
    e = [four.obs()["vnb"].sum()]
    for step in range(2, 6):
        four.run(1, step - 1)
        e.append(four.obs()["vnb"].sum())
    assert e[1] == e[0] and e[2] == e[0] and e[3] == e[0] and e[4] != e[0]
