"""ibx_step_noalloc (the step loop with the library's static scratch array for the
reduction partials and counters) against ibx_step, bit for bit: two calls in a
row on two systems, the second smaller, so the array holds the first call's
contents when the second starts; and calls mixed with ibx_step on one context,
with a multiple time step whose non-evaluating steps read the partials of the
previous call's last evaluation."""
import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _relaxed(sy):
    """The system with the state of a short on-device equilibration (a raw random-walk melt blows up)."""
    e = Engine(sy)
    workloads.equilibrate(e, sy, stage_steps=50, final_steps=100)
    s, v = e.state()
    e.close()
    sy.s, sy.v = s[:sy.natoms].copy(), units.vel_from_reduced(v)
    return sy


def _same(a, b):
    sa, va = a.state()
    sb, vb = b.state()
    assert np.all(np.isfinite(sa)) and np.all(np.isfinite(va))
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)
    oa, ob = a.obs(), b.obs()
    for k in ("vnb", "vbond", "vangle", "vtors", "pt", "ek"):
        assert np.array_equal(oa[k], ob[k]), k
    assert np.array_equal(a.averages(), b.averages())


def test_noalloc_equals_step_two_calls_second_smaller():
    big = _relaxed(system.cg_melt(nchains=40, nbeads=50, seed=31, avoid=0.8, dt=0.005))
    small = _relaxed(system.cg_melt(nchains=12, nbeads=50, seed=32, avoid=0.8, dt=0.005))
    for sy, n in ((big, 23), (small, 17)):
        print("system of", sy.natoms, "beads,", n, "steps", flush=True)
        a, b = Engine(sy), Engine(sy)
        a.step(n)
        b.step_noalloc(n)
        _same(a, b)
        a.close()
        b.close()


def test_noalloc_mixed_with_step_and_respa():
    sy = _relaxed(system.ua_melt(nchains=10, nbeads=100, seed=33, avoid=0.6, dt=0.001))
    a, b = Engine(sy), Engine(sy)
    for e in (a, b):
        e.set_respa(3)
    a.step(31)
    b.step(7)
    b.step_noalloc(11)       # starts on step 8, which reads step 7's nonbonded partials
    b.step(6)
    b.step_noalloc(7)
    print("counters", b.counters(), flush=True)
    _same(a, b)
