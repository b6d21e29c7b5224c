"""Error paths of the device path, each forced from the host so every kernel exits normally:

- a torsion, and an out-of-plane angle, outside its table (TORSIONS.inc:103-109,
  OUTOFPLANES.inc:93-100): IBX_E_TABLE, where the oracle also stops;
- bonds stretched past their table (BONDS.inc:21-31): counted, clamped to the
  last row, forces as the oracle's;
- a list that overflows during stepping (NEW_NEIGHBOUR_WITHLIST.f90:72, "LIST
  TOO SMALL"): IBX_E_LIST_OVERFLOW from ibx_step (the list build writes only
  within each row's capacity and the force kernel reads only that much), and
  the device keeps working for the next context;
- a served job larger than the serving loop's budget: failed with IBX_E_NOMEM
  while a job that fits completes.
"""
import copy

import numpy as np
import pytest

from ibisco_amd import system, units
from ibisco_amd.engine import Engine, IbxError, Server, estimate_bytes

pytestmark = pytest.mark.gpu

IBX_E_LIST_OVERFLOW, IBX_E_TABLE, IBX_E_NOMEM = -2, -3, -5


def test_torsion_outside_table(oracle_mod):
    s = system.ua_melt(nchains=20, nbeads=100, seed=5, avoid=0.6)
    t = s.torsion_tables[0]
    keep = np.asarray(t.x) <= 180.0                 # half the dihedral range
    s.torsion_tables = [system.Table(t.types, np.asarray(t.x)[keep], np.asarray(t.v)[keep], t.name)]
    o = oracle_mod.Oracle(s)
    with pytest.raises(RuntimeError):
        o.run(1, -1)
    e = Engine(s)
    with pytest.raises(IbxError) as ei:
        e.compute_forces(rebuild=True)
    print(ei.value, flush=True)
    assert ei.value.code == IBX_E_TABLE and "torsion" in str(ei.value)


def test_oop_outside_table(oracle_mod):
    s = system.comb_melt(nchains=8, seed=1)
    t = s.oop_tables[0]
    n = max(4, len(t.x) // 8)                       # the first eighth of the range only
    s.oop_tables = [system.Table(t.types, np.asarray(t.x)[:n], np.asarray(t.v)[:n], t.name)]
    e = Engine(s)
    with pytest.raises(IbxError) as ei:
        e.compute_forces(rebuild=True)
    print(ei.value, flush=True)
    assert ei.value.code == IBX_E_TABLE and "out-of-plane" in str(ei.value)


def test_bond_overextension_counted_and_clamped(oracle_mod):
    s = system.cg_melt(nchains=20, nbeads=50, seed=3, avoid=0.8)
    s.s = s.s.copy()
    for c in range(3):                              # the last bead of three chains 1.5 nm out (table ends 1.175)
        last = (c + 1) * 50 - 1
        s.s[last] = s.s[last - 1] + np.array([1.5, 0.0, 0.0])
    o = oracle_mod.Oracle(s)
    o.new_force()
    e = Engine(s)
    e.compute_forces(rebuild=True)
    og, oo = e.obs(), o.obs()
    rr = float(np.linalg.norm(e.forces() - o.forces()) / np.linalg.norm(o.forces()))
    print("over-extended bonds: device", og["overext"], "oracle", oo["overext"], "force rel RMS", rr, flush=True)
    assert oo["overext"] == 3 and og["overext"] == oo["overext"]
    assert rr <= 1e-5
    assert abs(og["vbond"].sum() - oo["vbond"].sum()) <= 1e-7 * abs(oo["vbond"].sum())


def test_list_overflow_during_stepping():
    """600 CG chains of 10 beads with every table zero (no forces: ballistic motion); the middle bead of each
    chain is aimed at the box centre so that all 600 meet exactly at the rebuild of step 10, far more
    neighbours than a row holds (1.5 x the mean count within rlist + 64)."""
    s = system.cg_melt(nchains=600, nbeads=10, seed=8)
    zero = lambda t: system.Table(t.types, t.x, np.zeros_like(np.asarray(t.v)), t.name)   # noqa: E731
    s.nb_tables = [zero(t) for t in s.nb_tables]
    s.bond_tables = [zero(t) for t in s.bond_tables]
    s.angle_tables = [zero(t) for t in s.angle_tables]
    dt = s.control.dt
    centre = np.zeros(3)
    mid = np.arange(600) * 10 + 5
    s.v = np.zeros_like(s.s)
    s.v[mid] = (centre - s.s[mid]) / (9 * dt)      # the rebuild of step 10 sees the positions after 9 moves
    e = Engine(s)
    e.step(9)                                      # no rebuild yet
    with pytest.raises(IbxError) as ei:
        e.step(1)
    print(ei.value, flush=True)
    assert ei.value.code == IBX_E_LIST_OVERFLOW and "LIST TOO SMALL" in str(ei.value)
    e.close()
    # the device is unharmed: a fresh context steps normally
    ok = Engine(system.cg_melt(nchains=10, nbeads=50, seed=2, avoid=0.8, dt=0.002))
    ok.step(10)
    assert np.all(np.isfinite(ok.state()[0]))


def test_served_job_over_budget():
    small = system.cg_melt(nchains=4, nbeads=25, seed=1, avoid=0.8)
    big = system.cg_melt(nchains=100, nbeads=50, seed=2, avoid=0.8)
    bs, bb = estimate_bytes(small, 150), estimate_bytes(big, 150)
    assert bs < bb
    srv = Server(hbm_budget=(bs + bb) // 2, max_cohorts=1, rdf_nbin=150, rdf_dr=0.01)
    jb = srv.submit(big, 20)
    js = srv.submit(small, 20)
    with pytest.raises(IbxError) as ei:
        srv.wait(jb)
    print(ei.value, flush=True)
    assert ei.value.code == IBX_E_NOMEM and "budget" in str(ei.value)
    r = srv.wait(js)
    assert r["nsteps"] == 20
    st = srv.stats()
    srv.close()
    assert st["failed"] == 1 and st["completed"] >= 1
