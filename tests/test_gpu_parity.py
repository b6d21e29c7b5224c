"""HIP path against the FP64 oracle on identical inputs (needs a gfx950 device).

Tolerances: pair sets and exclusions bit-exact; per-item
forces within 1e-5 relative RMS; energies and virial within 1e-7 relative to
the magnitude of their totals.
"""
import numpy as np
import pytest

from ibisco_amd import system
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-7


def _sys_cg(n_chains=40, seed=3, **kw):
    return system.cg_melt(nchains=n_chains, nbeads=50, seed=seed, avoid=0.8, **kw)


def _sys_ua(seed=5):
    return system.ua_melt(nchains=20, nbeads=100, seed=seed, avoid=0.6)


CASES = {
    "cg2000": lambda: _sys_cg(),
    "ua2000_torsion": lambda: _sys_ua(),
    "cg600_two_cells": lambda: system.cg_melt(nchains=12, nbeads=50, seed=7, avoid=0.8),
}


def _rel_rms(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check_energy(name, g, o, scale):
    err = np.abs(np.asarray(g) - np.asarray(o))
    assert np.all(err <= ENERGY_RTOL * scale), f"{name}: gpu={g} oracle={o} scale={scale}"


@pytest.mark.parametrize("case", sorted(CASES))
def test_pairs_bit_exact(case, oracle_mod):
    s = CASES[case]()
    o = oracle_mod.Oracle(s)
    e = Engine(s)
    po, pg = o.pairs(0), e.pairs(0, 0)
    print(case, "pairs", len(po), len(pg))
    assert len(po) == len(pg)
    assert np.array_equal(po, pg)


@pytest.mark.parametrize("case", sorted(CASES))
def test_forces_energy_virial(case, oracle_mod):
    s = CASES[case]()
    o = oracle_mod.Oracle(s)
    o.new_force()
    e = Engine(s)
    e.compute_forces(rebuild=True)
    fo, fg = o.forces(), e.forces()
    oo, og = o.obs(), e.obs()
    rr = _rel_rms(fg, fo)
    print(case, "force rel rms", rr)
    for k in ("vnb", "vbond", "vangle", "vtors", "pt"):
        print(case, k, og[k], oo[k])
    assert rr <= FORCE_RTOL
    epot = sum(float(np.abs(oo[k]).sum()) for k in ("vnb", "vbond", "vangle", "vtors", "voop"))
    for k in ("vnb", "vbond", "vangle", "vtors", "voop"):
        _check_energy(k, og[k], oo[k], max(float(np.abs(oo[k]).sum()), 1e-3 * epot))
    _check_energy("virial", og["pt"], oo["pt"], float(np.abs(oo["pt"][:3]).sum()))


def test_trajectory_short_nve(oracle_mod):
    s = _sys_cg(dt=0.001)
    o = oracle_mod.Oracle(s)
    e = Engine(s)
    o.run(20)
    e.step(20)
    so, vo, _ = o.state()
    sg, vg = e.state()
    ob, eb = o.obs(), e.obs()
    print("pos max abs", np.abs(sg - so).max(), "vel rel", _rel_rms(vg, vo), "ek", eb["ek"], ob["ek"])
    assert np.abs(sg - so).max() < 1e-6
    assert _rel_rms(vg, vo) < 1e-5
    assert abs(eb["ek"][0] - ob["ek"][0]) <= 1e-6 * abs(ob["ek"][0])


def test_batch_equals_alone():
    systems = [_sys_cg(n_chains=24, seed=11), _sys_cg(n_chains=40, seed=12), _sys_cg(n_chains=31, seed=13)]
    for k, sy in enumerate(systems):
        r, v = system.corrected_cg_table(k)
        sy.nb_tables[0] = system.Table((0, 0), r, v, "nb1")
        sy.control.dt = 0.002
    batch = Engine(systems)
    batch.step(25)
    for k, sy in enumerate(systems):
        alone = Engine(sy)
        alone.step(25)
        sa, va = alone.state()
        sb, vb = batch.state(k)
        assert np.array_equal(sa, sb), k
        assert np.array_equal(va, vb), k
        oa, ob = alone.obs(), batch.obs(k)
        assert np.array_equal(oa["vnb"], ob["vnb"]) and np.array_equal(oa["pt"], ob["pt"]), k


@pytest.mark.parametrize("calls", [(35,), (64, 64), (139,)])
def test_graph_equals_eager(calls):
    """Graph replays (blocks of one list interval and of four, eager head and tail) against eager steps:
    35 steps use the one-interval block only; 64 + 64 and 139 use both, the second call reusing them."""
    import copy
    from ibisco_amd import units, workloads
    s = _sys_cg(dt=0.002)
    pre = Engine(s)                     # equilibrated start: the raw melt heats up within ~60 steps
    workloads.equilibrate(pre, s, stage_steps=100, final_steps=200)
    t = copy.deepcopy(s)
    t.s, v = pre.state()
    t.v = units.vel_from_reduced(v)
    pre.close()
    a, b = Engine(t), Engine(t)
    a.step(sum(calls))
    for n in calls:
        b.step(n, graph=True)
    c = b.counters()
    print("counters", c, flush=True)
    assert c["graph_steps"] >= sum(calls) - len(calls) * (c["period"] - 1)
    assert c["graph_launches"] >= len(calls) and c["captures"] >= 1
    sa, va = a.state()
    sb, vb = b.state()
    assert np.array_equal(sa, sb)
    assert np.array_equal(va, vb)
    assert np.array_equal(a.obs()["pt"], b.obs()["pt"])


def test_one_force_launch_equals_separate_launches():
    """The step loop's single force launch (nonbonded, bonded and the previous
    step's field sums as roles of one kernel) against the same step with each
    kernel class launched on its own (the timed path): the same bits."""
    s = _sys_cg(dt=0.002)
    a, b = Engine(s), Engine(s)
    b.set_timing(True)
    for e in (a, b):
        e.step(13)
        e.step(22, graph=e is a)
    sa, va = a.state()
    sb, vb = b.state()
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)
    oa, ob = a.obs(), b.obs()
    for k in ("vnb", "vbond", "vangle", "vtors", "pt"):
        assert np.array_equal(oa[k], ob[k]), k
    assert np.array_equal(a.averages(), b.averages())


def test_back_run_after_large_displacement(oracle_mod):
    """The force kernel skips the back run of each list (pairs beyond rc + delta at the rebuild) only while
    every member has moved less than delta/2.  An equilibrated melt is run 30 steps without a rebuild, so
    members move further and back-run pairs come inside rc: forces at the final positions must equal those
    of a fresh list (the oracle), which holds as long as no pair beyond rlist came inside rc."""
    import copy
    from ibisco_amd import units, workloads
    s = _sys_cg(n_chains=30, seed=21)
    pre = Engine(s)
    workloads.equilibrate(pre, s, stage_steps=100, final_steps=400)
    s_eq, v_eq = pre.state()
    pre.close()
    t = copy.deepcopy(s)
    t.s, t.v = s_eq.copy(), units.vel_from_reduced(v_eq)
    t.control.nupdate = 10 ** 6
    e = Engine(t)
    e.step(30)
    e.compute_forces(rebuild=False)
    sg, vg = e.state()
    fg = e.forces()
    delta = (t.control.rlist - t.control.cutoff) / 3.0
    moved = float(np.sqrt(((sg - s_eq) ** 2).sum(1)).max())
    print("largest displacement", moved, "delta/2", delta / 2, flush=True)
    assert delta / 2 < moved < 0.5 * (t.control.rlist - t.control.cutoff)
    o = oracle_mod.Oracle(t)
    o.set_state(sg, vg)
    o.update_lists()
    o.new_force()
    rr = _rel_rms(fg, o.forces())
    print("force rel rms after", moved, "nm:", rr)
    assert rr <= FORCE_RTOL


def test_crowded_cell(oracle_mod):
    """A cell holding more members than the cell sort ranks in LDS (64): 200 of 2,000 single beads placed within
    0.01 nm of one list cell's centre (cells of rlist/3 along x and y, rlist/6 along z) are ranked in global
    memory.  The pair set is bit-exact against the oracle, and the nonbonded energy, virial and forces of the
    system batched behind another equal its own alone bit for bit (the slot order is deterministic)."""
    s = system.cg_melt(nchains=2000, nbeads=1, seed=21, avoid=0.8)
    rl, L = s.control.rlist, np.asarray(s.box, float)
    nc = np.array([int(3 * L[0] / rl), int(3 * L[1] / rl), int(6 * L[2] / rl)])
    centre = ((nc // 2) + 0.5) / nc * L - 0.5 * L
    rng = np.random.default_rng(4)
    s.s[:200] = centre + rng.uniform(-0.01, 0.01, (200, 3))
    folded = s.s - L * np.round(s.s / L)
    cell = np.clip(((folded / L + 0.5) * nc).astype(int), 0, nc - 1)
    _, counts = np.unique(cell[:, 0] * nc[1] * nc[2] + cell[:, 1] * nc[2] + cell[:, 2], return_counts=True)
    assert counts.max() >= 200, counts.max()
    o = oracle_mod.Oracle(s)
    e = Engine(s)
    po, pg = o.pairs(0), e.pairs(0, 0)
    print("crowded cell: pairs", len(po), len(pg), "largest cell", counts.max(), flush=True)
    assert np.array_equal(po, pg)
    e.compute_forces(rebuild=True)
    other = system.cg_melt(nchains=500, nbeads=1, seed=22, avoid=0.8)
    batch = Engine([other, s])
    batch.compute_forces(rebuild=True)
    oa, ob = e.obs(), batch.obs(1)
    assert np.array_equal(oa["vnb"], ob["vnb"]) and np.array_equal(oa["pt"], ob["pt"])
    assert np.array_equal(e.forces(), batch.forces(1))
