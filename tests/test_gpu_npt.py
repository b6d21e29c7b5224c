"""The NPT barostat (SCALEBP.f90:6-44) on the device against the FP64 oracle (needs a gfx950 device).

The oracle's barostat is pinned to the reference's own NPT run of hybridPA in
test_golden.py.  Here:
- hybridPA in NPT (the testsuite's hybridPA-NPT input: virtual sites, two
  neighbour groups, list rebuilt every 20 steps): box, positions and velocities
  after 10 and 40 steps against the oracle, and the box against the reference's
  run;
- a batch of three coarse-grained melts from one equilibrated state: one with a
  target pressure far above its own (its box shrinks by 1 to 2 % at each
  rescaling, and its cells start 0.3 % above rlist / 3, so the list build's
  cell grid must follow the box down), one far below (the box grows past the
  grid it was sized for), one NVT beside them; each against the oracle, and the
  lists rebuilt at the final boxes against the oracle's at the same positions;
- graph replay against eager steps, bit for bit;
- an input directory in NPT against the same system given in memory.
"""
import copy
import os

import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX_RTOL_ORACLE = 1e-9    # device vs oracle: MIU from FP32 forces' virial, FP64 otherwise (measured 3e-11)
BOX_RTOL_REF = 1e-6       # device vs the reference's REAL*4 run (the oracle's own gate; measured 7e-8)
POS_ATOL = 1e-7           # nm after 40 steps (measured 1.4e-8 hybridPA, 1.6e-9 CG melts)
VEL_RTOL = 1e-6           # relative RMS after 40 steps (measured 7e-8)


def _rel_rms(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_hybrid_npt_matches_oracle_and_reference(oracle_mod):
    z = np.load(os.path.join(HERE, "hybridPA_npt.npz"))
    sysm = system.from_arrays({k[4:]: z[k] for k in z.files if k.startswith("sys_")})
    assert sysm.control.ensemble == "NPT"
    print("case hybridPA_npt", flush=True)
    o = oracle_mod.Oracle(sysm)
    e = Engine(sysm)
    box0 = np.asarray(sysm.box, np.float64)
    done = 0
    na = sysm.natoms
    for n in (10, 40):
        o.run(n - done, done)
        e.step(n - done)
        done = n
        so, vo, _ = o.state()
        sg, vg = e.state()
        bo, bg, br = o.obs()["box"], e.box(), z[f"run{n}_box"]
        dbox = np.abs(bg - bo).max() / bo.max()
        dref = np.abs(bg - br).max() / br.max()
        dpos = np.abs(sg[:na] - so[:na]).max()
        rv = _rel_rms(vg, vo)
        print(f"hybridPA NPT {n} steps: box {bg[0]:.9f} (oracle {bo[0]:.9f}, reference {br[0]:.9f}) "
              f"rel {dbox:.2e} / {dref:.2e}; atoms max |ds| {dpos:.2e}, velocity rel rms {rv:.2e}", flush=True)
        assert np.abs(bg - box0).max() > 1e-3 * box0.max()   # the box did move
        assert dbox <= BOX_RTOL_ORACLE and dref <= BOX_RTOL_REF
        assert dpos <= POS_ATOL and rv <= VEL_RTOL
        assert abs(e.obs()["volume"] - np.prod(bg)) <= 1e-12 * np.prod(bg)


@pytest.fixture(scope="module")
def melt_state():
    s = system.cg_melt(nchains=40, nbeads=50, seed=3, avoid=0.8, dt=0.002)
    pre = Engine(s)
    workloads.equilibrate(pre, s, stage_steps=100, final_steps=200)
    t = copy.deepcopy(s)
    t.s, v = pre.state()
    t.v = units.vel_from_reduced(v)
    pre.close()
    return t


def _npt_batch(s, oracle_mod):
    """Compressed, expanded and NVT copies of s: the targets sit 15 GPa above and below the mean pressure of the
    first ten steps (what the first rescaling averages), so each rescaling changes the box edge by 1 to 2 %."""
    t = copy.deepcopy(s)
    t.control.ensemble = "NVT"
    o = oracle_mod.Oracle(t)
    ps = []
    for k in range(10):
        o.run(1, k)
        ps.append(float(o.obs()["pt"][:3].mean()))
    pbar = float(np.mean(ps)) * units.PSCALE
    out = []
    for ens, dp, rl in (("NPT", 1.5e7, 3 * s.box[0] / 10.03), ("NPT", -1.5e7, 1.8), ("NVT", 0.0, 1.8)):
        t = copy.deepcopy(s)
        t.control.ensemble = ens
        t.control.pressure = pbar + dp
        t.control.rlist = t.control.bead_rlist = rl
        out.append(t)
    return out


def test_npt_batch_matches_oracle(melt_state, oracle_mod):
    systems = _npt_batch(melt_state, oracle_mod)
    e = Engine(systems)
    e.step(40)
    for k, t in enumerate(systems):
        o = oracle_mod.Oracle(t)
        o.run(40)
        so, vo, _ = o.state()
        sg, vg = e.state(k)
        bo, bg = o.obs()["box"], e.box(k)
        dbox = np.abs(bg - bo).max() / bo.max()
        dpos = np.abs(sg - so).max()
        rv = _rel_rms(vg, vo)
        print(f"system {k} ({t.control.ensemble}): box {bg[0]:.6f} from {t.box[0]:.6f}, cells per edge "
              f"{3 * bg[0] / t.control.rlist:.3f}; rel box {dbox:.2e}, max |ds| {dpos:.2e}, vel rel rms {rv:.2e}",
              flush=True)
        assert dbox <= BOX_RTOL_ORACLE
        assert dpos <= POS_ATOL and rv <= VEL_RTOL
        if k == 0:
            assert 3 * bg[0] / t.control.rlist < 9.5            # the grid went from 10 cells per edge to 9
        elif k == 1:
            assert bg[0] > 1.02 * t.box[0]
        else:
            assert np.array_equal(bg, np.asarray(t.box, np.float64))
    # the lists at the final boxes: the same pairs as the oracle's list build at the device's positions and box
    e.rebuild_lists()
    for k, t in enumerate(systems):
        sg, vg = e.state(k)
        u = copy.deepcopy(t)
        u.s = sg[:t.natoms].copy()
        u.v = units.vel_from_reduced(vg)
        u.box = e.box(k)
        o = oracle_mod.Oracle(u)
        o.new_force()
        po, pg = o.pairs(0), e.pairs(k, 0)
        print(f"system {k}: {len(po)} pairs (oracle), {len(pg)} (device)", flush=True)
        assert len(po) > 0 and np.array_equal(po, pg)


def test_npt_graph_equals_eager(melt_state, oracle_mod):
    systems = _npt_batch(melt_state, oracle_mod)
    a, b = Engine(systems), Engine(systems)
    a.step(70)
    b.step(70, graph=True)
    assert b.counters()["graph_steps"] > 0
    for k in range(len(systems)):
        sa, va = a.state(k)
        sb, vb = b.state(k)
        assert np.array_equal(sa, sb) and np.array_equal(va, vb), k
        assert np.array_equal(a.box(k), b.box(k)), k


def test_npt_input_directory_equals_in_memory(tmp_path, melt_state, oracle_mod):
    """RDCONTROL.f90's pressure, pressure_coupling_time and isothermal_compressibility read from a control file
    (ibx_load_dir) give the barostat that the same files read in Python and passed in memory give: the same
    box and trajectory, up to the last bits of the unit conversions done on the two sides."""
    t = _npt_batch(melt_state, oracle_mod)[0]
    d = str(tmp_path)
    system.write_dir(t, d)
    a = Engine(system.read_dir(d))
    b = Engine(None, load_dir=d)
    a.step(30)
    b.step(30)
    sa, va = a.state()
    sb, vb = b.state()
    ba, bb = a.box(), b.box()
    print("input directory: box", bb, "in memory", ba, "start", t.box, flush=True)
    assert ba[0] < 0.98 * t.box[0]
    assert np.abs(ba - bb).max() <= 1e-12 * ba.max()
    assert np.abs(sa - sb).max() <= 1e-9 and _rel_rms(vb, va) <= 1e-9
