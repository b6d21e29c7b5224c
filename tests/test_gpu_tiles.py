"""Layouts of the force launch and the tile autotuner (ibx_set_tiles; no
counterpart in the reference).

One workgroup per block, and persistent grids taking blocks from per-XCD work
queues, compute every block the same way, so states, observables and window
sums after a run must be the same bits whichever layout ran.  The cases cover
cubic and orthorhombic boxes, the united-atom density (torsions), a hybrid
system (two groups, virtual sites) and a batch of replicas of different sizes,
each with its own table.
"""
import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

LAYOUTS = [(0, -1, -1), (1, 1, 5), (1, 2, 5), (1, 1, 4), (1, 3, 1)]


def _cases():
    cg = system.cg_melt(nchains=60, nbeads=50, seed=21, avoid=0.8, dt=0.002)
    ortho = system.cg_melt(nchains=50, nbeads=50, seed=22, avoid=0.8, dt=0.002, box=(7.0, 5.6, 8.75))
    ua = system.ua_melt(nchains=40, nbeads=100, seed=23, avoid=0.6)
    hyb = system.hybrid_melt(20, 47, chain_len=60, seed=24, avoid=0.8, light=True)
    hyb.control.dt = 0.0005
    reps = []
    for k, (nc, nb) in enumerate([(30, 40), (90, 60), (12, 25)]):
        s = system.cg_melt(nchains=nc, nbeads=nb, seed=30 + k, avoid=0.8, dt=0.002)
        r, v = system.corrected_cg_table(k)
        s.nb_tables[0] = system.Table((0, 0), r, v, "nb1")
        reps.append(s)
    return {"cg": [cg], "ortho": [ortho], "ua": [ua], "hybrid": [hyb], "replicas": reps}


CASES = _cases()


def _equilibrated(systems):
    """Random-walk starts are far from equilibrium (some would not survive a run): equilibrate first."""
    e = Engine(systems)
    workloads.equilibrate(e, systems, stage_steps=50, final_steps=200)
    for k, s in enumerate(systems):
        pos, vel = e.state(k)
        s.s, s.v = pos[:s.natoms].copy(), units.vel_from_reduced(vel)
    e.close()
    return systems


@pytest.mark.parametrize("name", sorted(CASES))
def test_layouts_give_the_same_bits(name):
    systems = _equilibrated(CASES[name])
    print("case", name, [s.natoms for s in systems], flush=True)
    ref = None
    for lay in LAYOUTS:
        e = Engine(systems)
        e.set_tiles(*lay)
        e.step(45, graph=True)   # graph blocks and an eager tail, four rebuilds
        t, _ = e.tiles()
        assert t[0] == lay[0] and (lay[0] == 0 or (t[1] == lay[1] and t[2] == lay[2])), (lay, t)
        out = []
        for k in range(len(systems)):
            sp, vp = e.state(k)
            ob = e.obs(k)
            out.append((sp, vp, ob["vnb"], ob["pt"], ob["ek"], e.averages(k)))
        e.close()
        assert all(np.all(np.isfinite(o[0])) for o in out)
        if ref is None:
            ref = out
            continue
        for k, (a, b) in enumerate(zip(ref, out)):
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (name, lay, k)


def test_autotuner_picks_a_timed_layout():
    systems = CASES["replicas"]
    e = Engine(systems)
    t, ms = e.tiles()
    assert t == [-1, -1, -1]
    e.step(12, graph=True)
    t, ms = e.tiles()
    print("tiles", t, "ms per candidate", ms, flush=True)
    assert all(m > 0 for m in ms)
    cand = [(0, 0, 0), (1, 1, 5), (1, 2, 5), (1, 1, 4)]
    assert tuple(t) in cand
    assert ms[cand.index(tuple(t))] == min(ms)
    e.set_tiles(0)             # a fixed layout: applied, not timed
    e.step(3)
    t, ms = e.tiles()
    assert t == [0, 0, 0] and all(m == -1 for m in ms)
    e.close()


def test_set_tiles_rejects_bad_layouts():
    e = Engine(CASES["cg"])
    for bad in [(2, -1, -1), (1, 0, 5), (1, 1, 0), (1, 1, 17), (-2, -1, -1)]:
        with pytest.raises(Exception):
            e.set_tiles(*bad)
    e.close()
