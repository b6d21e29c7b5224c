"""The list-skin autotuner (no counterpart in the reference).

It moves only the boundary between the two runs of each list row, so the pair
set must stay the oracle's bit for bit, forces must stay within the usual
gates of the oracle's, and a tuned run may differ from an untuned one only by
FP32 summation order.
"""
import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _sys(seed=3):
    """An equilibrated CG melt (random-walk starts are far from equilibrium)."""
    s = system.cg_melt(nchains=40, nbeads=50, seed=seed, avoid=0.8, dt=0.002)
    e = Engine(s)
    workloads.equilibrate(e, s, stage_steps=50, final_steps=200)
    pos, vel = e.state()
    e.close()
    s.s, s.v = pos[:s.natoms].copy(), units.vel_from_reduced(vel)
    return s


def test_tuned_skin_keeps_pairs_and_forces(oracle_mod):
    s = _sys()
    skin = s.control.rlist - s.control.cutoff
    e = Engine(s)
    assert abs(e.skin(0)[0] - skin / 3) < 1e-5
    e.step(30, graph=True)             # tuned at the rebuilds of steps 10, 20 and 30
    d = e.skin(0)[0]
    print("tuned delta", d, "of skin", skin, flush=True)
    assert skin / 16 - 1e-6 <= d <= skin + 1e-6 and abs(d - skin / 3) > 1e-4
    e.rebuild_lists()                  # a list at the current positions with the tuned delta
    pos, vel = e.state()
    o = oracle_mod.Oracle(s)
    o.set_state(pos[:s.natoms], vel)
    o.update_lists()
    assert np.array_equal(o.pairs(0), e.pairs(0, 0))
    o.new_force()
    e.compute_forces(rebuild=False)
    fo, fg = o.forces(), e.forces()
    rr = float(np.linalg.norm(fg - fo) / np.linalg.norm(fo))
    print("force rel rms", rr, flush=True)
    assert rr <= 1e-5


def test_tuned_and_untuned_runs_agree():
    s = _sys(seed=5)
    a, b = Engine(s), Engine(s)
    b.set_skin_tuning(False)
    a.step(40, graph=True)
    b.step(40, graph=True)
    sa, va = a.state()
    sb, vb = b.state()
    print("tuned", a.skin(0)[0], "untuned", b.skin(0)[0], "max |ds|", np.abs(sa - sb).max(), flush=True)
    assert np.all(np.isfinite(sa)) and np.all(np.isfinite(sb))
    assert abs(b.skin(0)[0] - (s.control.rlist - s.control.cutoff) / 3) < 1e-5
    assert np.abs(sa - sb).max() <= 1e-6
    assert float(np.linalg.norm(va - vb) / np.linalg.norm(vb)) <= 1e-5
