"""Hybrid systems with centre-of-mass virtual sites: HIP path against the oracle.

Atomistic chains (two atom masses) carry one COM site per 6 atoms; CG bead
chains interact with the sites through their own table (bead-site), atoms with
atoms through the atom table.  The oracle's COM sites are pinned in
test_hybrid_com.py.  Gates as in test_gpu_parity.py: both groups' pair sets
bit-exact, forces within 1e-5 relative RMS, energies and virial within 1e-7,
and a short NVE run from an equilibrated state (site positions rebuilt from
the moved atoms every step, site forces spread by mass fraction).
"""
import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-7


def _hybrid(seed=2):
    return system.hybrid_melt(12, 40, chain_len=60, seed=seed, avoid=0.8, light=True)


def _rel_rms(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_hybrid_com_pairs_forces_energy(oracle_mod):
    s = _hybrid()
    print("hybrid", s.natoms, "atoms+beads", s.nvs, "sites", flush=True)
    o = oracle_mod.Oracle(s)
    o.new_force()
    e = Engine(s)
    for g in (0, 1):
        po, pg = o.pairs(g), e.pairs(0, g)
        print("group", g, "pairs", len(po), len(pg), flush=True)
        assert len(po) > 0 and np.array_equal(po, pg)
    e.compute_forces(rebuild=True)
    fo, fg = o.forces(), e.forces()
    rr = _rel_rms(fg, fo)
    print("force rel rms", rr, flush=True)
    assert rr <= FORCE_RTOL
    oo, og = o.obs(), e.obs()
    epot = sum(float(np.abs(oo[k]).sum()) for k in ("vnb", "vbond", "vangle"))
    for k in ("vnb", "vbond", "vangle"):
        scale = max(float(np.abs(oo[k]).sum()), 1e-3 * epot)
        assert np.all(np.abs(og[k] - oo[k]) <= ENERGY_RTOL * scale), (k, og[k], oo[k])
    assert np.all(np.abs(og["pt"] - oo["pt"]) <= ENERGY_RTOL * float(np.abs(oo["pt"][:3]).sum()))
    # site positions on the device are the oracle's centres of mass
    sg, _ = e.state()
    so, _, _ = o.state()
    assert np.abs(sg[s.natoms:] - so[s.natoms:]).max() <= 1e-12 * np.abs(so).max()


def test_hybrid_com_short_nve_run(oracle_mod):
    s = _hybrid(seed=3)
    s.control.dt = 0.0005
    e = Engine(s)
    workloads.equilibrate(e, s, stage_steps=50, final_steps=200)
    pos, vel = e.state()
    e.close()
    s.s, s.v = pos[:s.natoms].copy(), units.vel_from_reduced(vel)
    o = oracle_mod.Oracle(s)
    e = Engine(s)
    o.run(20)
    e.step(20)
    so, vo, _ = o.state()
    sg, vg = e.state()
    na = s.natoms
    dpos = np.abs(sg[:na] - so[:na]).max()
    rv = _rel_rms(vg, vo)
    # the device places the sites after MOVE (the reference and the oracle at the start of the next
    # step, VIRTUAL_DEF.f90 in NEW_LOOP.F90's order): the returned sites are the centres of the returned atoms
    m = s.type_mass[s.itype]
    com = np.array([(m[a, None] * sg[a]).sum(0) / m[a].sum() for a in s.vs_atoms])
    print("after 20 steps: atoms max |ds|", dpos, "velocity rel rms", rv, "sites vs centres",
          np.abs(sg[na:] - com).max(), flush=True)
    assert dpos <= 1e-6 and rv <= 1e-5
    assert np.abs(sg[na:] - com).max() <= 1e-12 * np.abs(sg).max()
    oo, og = o.obs(), e.obs()
    et_o = sum(float(oo[k].sum()) for k in ("vnb", "vbond", "vangle")) + float(oo["ek"].sum())
    et_g = sum(float(og[k].sum()) for k in ("vnb", "vbond", "vangle")) + float(og["ek"].sum())
    assert abs(et_g - et_o) <= 1e-6 * abs(et_o)
