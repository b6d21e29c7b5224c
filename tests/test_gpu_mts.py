"""The MTS cycle of MTS_LOOP.f90:71-224 on the device (hybridPA, NMTS 2, NVE).

The oracle's cycle is pinned to the reference's MTS_FORCE by the hybridPA_mts
fixture (test_golden.py).  Here the device runs the same 12 steps: per-step
energies within the usual gates of the oracle's (the bead group's energy and
virial are extrapolated from saved partial sums, the oracle extrapolates the
totals), the reference's own per-step energies within the oracle's tolerance,
the final state as in the other short runs, and graph replay equal to eager
launches bit for bit.
"""
import os

import numpy as np
import pytest

from ibisco_amd import system
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ENERGY_RTOL = 1e-7


def _load():
    z = np.load(os.path.join(HERE, "hybridPA_mts.npz"))
    return z, system.from_arrays({k[4:]: z[k] for k in z.files if k.startswith("sys_")})


def _energies(ob):
    return np.concatenate([ob[k] for k in ("vnb", "vbond", "vangle", "vtors", "voop")])


def test_mts_cycle_matches_oracle_and_reference(oracle_mod):
    z, s = _load()
    assert s.control.mts == 2
    o = oracle_mod.Oracle(s, mts_quirk=True)
    e = Engine(s, mts_quirk=True)
    ref = z["run12_obs"]
    for k in range(12):
        o.run(1, k)
        e.step(1)
        eo, eg = _energies(o.obs()), _energies(e.obs())
        scale = np.abs(eo).sum()
        print("step", k + 1, "V_NB gpu", eg[:3].sum(), "oracle", eo[:3].sum(), "reference", ref[k, :3].sum(),
              flush=True)
        assert np.all(np.abs(eg - eo) <= ENERGY_RTOL * scale), (k + 1, eg, eo)
        assert np.all(np.abs(eg - ref[k, :15]) <= 1e-4 * np.abs(ref[k, :15]).sum())
        pto, ptg = o.obs()["pt"], e.obs()["pt"]
        assert np.all(np.abs(ptg - pto) <= ENERGY_RTOL * np.abs(pto[:3]).sum()), (k + 1, ptg, pto)
    so, vo, _ = o.state()
    sg, vg = e.state()
    na = s.natoms
    print("after 12 steps: atoms max |ds|", np.abs(sg[:na] - so[:na]).max(), flush=True)
    assert np.abs(sg[:na] - so[:na]).max() <= 1e-6
    assert float(np.linalg.norm(vg - vo) / np.linalg.norm(vo)) <= 1e-5


def test_mts_graph_equals_eager():
    """Graph blocks of whole MTS cycles, from a step inside a cycle (7) and again (25 steps: two blocks of the
    10-step period from phase 7, then eager steps), against eager launches: the same bits, with graph
    replays actually taken."""
    _, s = _load()
    a, b = Engine(s, mts_quirk=True), Engine(s, mts_quirk=True)
    a.step(32)
    b.step(7, graph=True)
    b.step(25, graph=True)
    c = b.counters()
    print("counters", c, flush=True)
    assert c["period"] == 10 and c["graph_launches"] >= 2 and c["graph_steps"] >= 20
    sa, va = a.state()
    sb, vb = b.state()
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)
    assert np.array_equal(a.obs()["vnb"], b.obs()["vnb"])
