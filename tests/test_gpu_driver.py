"""The Fortran step driver (ibisco_amd/fortran, iso_c_binding over libibx.so) on the GPU.

It reads an input directory in the reference's formats through ibx_load_dir
and steps it in blocks; its printed energies must equal those of the Python
binding driving the same library on the same directory, and the potential
energy after the first block must match the FP64 oracle's.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ibisco_amd import system, units
from ibisco_amd.build import DRIVER
from ibisco_amd.engine import Engine
from ibisco_amd.units import CONV

pytestmark = pytest.mark.gpu


def _run_driver(d, nsteps, block, *extra):
    assert os.path.exists(DRIVER), "ibx_driver is not built: run `python -m ibisco_amd.build`"
    r = subprocess.run([DRIVER, d, str(nsteps), str(block), *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [list(map(float, ln.split())) for ln in r.stdout.splitlines() if ln.strip() and not ln.startswith("#")]
    return np.array(rows)


def test_fortran_driver_matches_python_binding(oracle_mod):
    sysm = system.cg_melt(nchains=20, nbeads=50, seed=4, avoid=0.8, dt=0.002)
    with tempfile.TemporaryDirectory() as d:
        system.write_dir(sysm, d)
        print("driver on", d, flush=True)
        rows = _run_driver(d, 20, 10)
        assert rows.shape == (2, 9) and list(rows[:, 0]) == [10.0, 20.0]
        e = Engine(None, load_dir=d)
        got = []
        for _ in range(2):
            e.step(10, graph=True)
            ob = e.obs()
            epot = sum(float(ob[k].sum()) for k in ("vnb", "vbond", "vangle", "vtors", "voop"))
            got.append(epot * CONV)
        e.close()
        print("driver Epot", rows[:, 2], "python Epot", got)
        assert np.allclose(rows[:, 2], got, rtol=1e-10, atol=0)
        # the first block against the FP64 oracle on the same input files
        o = oracle_mod.Oracle(system.read_dir(d))
        o.run(10)
        oo = o.obs()
        eo = sum(float(oo[k].sum()) for k in ("vnb", "vbond", "vangle", "vtors", "voop")) * CONV
        assert abs(rows[0, 2] - eo) <= 1e-6 * abs(eo)


def test_fortran_driver_runs_the_mts_cycle():
    """hybridPA with `mts 2` in its control file: the driver's blocks end on the steps that take the
    reference's backward extrapolation (MOD(step, 5) = 0); their nonbonded energy must match the
    reference's own MTS_FORCE run of the same input (fixture hybridPA_mts)."""
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hybridPA_mts.npz"))
    sysm = system.from_arrays({k[4:]: z[k] for k in z.files if k.startswith("sys_")})
    ref = z["run12_obs"]
    with tempfile.TemporaryDirectory() as d:
        system.write_dir(sysm, d)
        print("driver on", d, flush=True)
        rows = _run_driver(d, 10, 5)
    assert rows.shape == (2, 9) and list(rows[:, 0]) == [5.0, 10.0]
    for k, step in enumerate((5, 10)):
        vnb_ref = ref[step - 1, 0:3].sum() * CONV
        epot_ref = ref[step - 1, 0:15].sum() * CONV
        print("step", step, "V_NB driver", rows[k, 5], "reference", vnb_ref, flush=True)
        assert abs(rows[k, 5] - vnb_ref) <= 1e-4 * abs(epot_ref)
        assert abs(rows[k, 2] - epot_ref) <= 1e-4 * abs(epot_ref)


def _records(b):
    out, off = [], 0
    while off < len(b):
        n = int(np.frombuffer(b, "<u4", 1, off)[0])
        out.append(b[off + 4:off + 4 + n])
        off += 8 + n
    return out


def test_fortran_driver_writes_the_trajectory():
    """With a trajectory file the driver appends a TRZ frame per block (WRITETRJ.F90 format): two frames after
    20 steps in blocks of 10, whose positions and velocities are the state the Python binding reaches, in single
    precision, and whose energies are its observables."""
    sysm = system.cg_melt(nchains=20, nbeads=50, seed=4, avoid=0.8, dt=0.002)
    with tempfile.TemporaryDirectory() as d:
        system.write_dir(sysm, d)
        trj = os.path.join(d, "s-md.trj")
        print("driver on", d, flush=True)
        _run_driver(d, 20, 10, trj)
        rec = _records(open(trj, "rb").read())
        e = Engine(None, load_dir=d)
        e.step(20, graph=True)
        s, v = e.state()
        ob = e.obs()
        e.close()
    n = sysm.natoms
    assert len(rec) == 2 + 2 * 10 and rec[0].decode().strip() == sysm.title
    assert np.frombuffer(rec[12], "<i4", 3).tolist() == [2, 20, n]
    for k in range(3):
        assert np.array_equal(np.frombuffer(rec[12 + 4 + k], "<f4"), s[:n, k].astype(np.float32))
        vk = (v[:, k] * (1.0 / 1.0e12 / units.TIMESCALE)).astype(np.float32)
        assert np.array_equal(np.frombuffer(rec[12 + 7 + k], "<f4"), vk)
    epot = sum(float(ob[k].sum()) for k in ("vnb", "vbond", "vangle", "vtors", "voop"))
    assert abs(np.frombuffer(rec[12 + 3], "<f8", 1, 12)[0] - epot * CONV) <= 1e-12 * abs(epot * CONV)
