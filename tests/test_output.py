"""Trajectory and restart writers against the files the reference's own WRITETRJ wrote (WRITETRJ.F90:10-178,
run by the reference harness after 10 steps of cg2000; tests/golden/trz_cg2000.npz).

Fed the values the reference wrote from, the TRZ writer must reproduce the record structure, header, integers,
box, temperature field and positions bit for bit; the fields the reference forms with its single-precision unit
constants (pressure, energies, velocities, time) to single precision.  The restart file must carry the same
numbers (positions to their 14 printed digits) and read back, as a coordinate file, into the same state.
"""
import os
import re
import shutil
import tempfile

import numpy as np
import pytest

from ibisco_amd import engine, system, units

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
Z = np.load(os.path.join(HERE, "trz_cg2000.npz"))


def _records(b):
    """Payloads of Fortran unformatted sequential records (4-byte length markers on both sides)."""
    out, off = [], 0
    while off < len(b):
        n = int(np.frombuffer(b, "<u4", 1, off)[0])
        out.append(b[off + 4:off + 4 + n])
        assert int(np.frombuffer(b, "<u4", 1, off + 4 + n)[0]) == n
        off += 8 + n
    return out


def _time_ps():
    return int(Z["nsteps"]) * float(Z["dt"]) * units.TIMESCALE * 1e12 + float(Z["initime"])


@pytest.fixture(scope="module")
def ours(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("trz") / "ours.trj")
    for frame in (1, 2):
        engine.trz_write_frame(p, str(Z["title"]), frame, int(Z["ntrj"]), _time_ps(), Z["box"], Z["pt"],
                               float(Z["epot"]), float(Z["ek"]), float(Z["mktemp"]), Z["s"].astype(np.float64),
                               Z["v"].astype(np.float64))
    return open(p, "rb").read()


def test_trz_layout_and_exact_fields(ours):
    ref = Z["trj"].tobytes()
    assert len(ours) == len(ref)
    ro, rr = _records(ours), _records(ref)
    assert [len(x) for x in ro] == [len(x) for x in rr]
    assert ro[0] == rr[0]                                   # title, 80 characters
    assert ro[1] == rr[1] == np.int32(10).tobytes()         # nrec
    natoms = int(Z["natoms"])
    for f in range(2):
        base = 2 + 10 * f
        assert ro[base][:12] == rr[base][:12]               # frame, frame * ntrj, atoms
        assert ro[base + 1] == rr[base + 1]                 # the box
        assert ro[base + 3][:4] == rr[base + 3][:4]         # 6
        eo = np.frombuffer(ro[base + 3], "<f8", 6, 4)
        er = np.frombuffer(rr[base + 3], "<f8", 6, 4)
        assert eo[3] == er[3] and eo[4] == er[4] == 0.0     # temperature field, zeros
        for k in range(3):
            assert ro[base + 4 + k] == rr[base + 4 + k]     # positions, single precision
            vo = np.frombuffer(ro[base + 7 + k], "<f4", natoms)
            vr = np.frombuffer(rr[base + 7 + k], "<f4", natoms)
            assert np.allclose(vo, vr, rtol=2e-6, atol=0)
        to = np.frombuffer(ro[base], "<f8", 1, 12)[0]
        tr = np.frombuffer(rr[base], "<f8", 1, 12)[0]
        assert abs(to - tr) <= 1e-6 * abs(tr)


def test_trz_single_precision_fields(ours):
    ro, rr = _records(ours), _records(Z["trj"].tobytes())
    po, pr = np.frombuffer(ro[4], "<f8"), np.frombuffer(rr[4], "<f8")   # pressure and tensor, kPa
    assert np.allclose(po, pr, rtol=2e-6, atol=1e-6 * np.abs(pr).max())
    eo = np.frombuffer(ro[5], "<f8", 3, 4)
    er = np.frombuffer(rr[5], "<f8", 3, 4)
    assert np.allclose(eo, er, rtol=2e-6, atol=0)


NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eEdD][-+]?\d+)?$")


def _numbers(text):
    """Every numeric token after the record description (the molecule and atom records)."""
    lines = text.splitlines()
    toks = " ".join(lines[9:]).split()
    return [float(t.replace("D", "E")) if NUM.match(t) else t for t in toks]


def test_restart_carries_the_reference_numbers(tmp_path):
    p = str(tmp_path / "restart")
    engine.restart_write(p, str(Z["title"]), _time_ps(), Z["box"], Z["itype"], Z["nbonds"], Z["jbond"],
                         Z["s"].astype(np.float64), Z["v"].astype(np.float64), Z["mol_sizes"], Z["mol_names"])
    ours = _numbers(open(p).read())
    ref = _numbers(Z["restart"].tobytes().decode())
    assert len(ours) == len(ref)
    for a, b in zip(ours, ref):
        if isinstance(b, str):
            assert a == b
        elif float(b).is_integer() and "." not in str(b) and abs(b) < 1e6 and isinstance(a, float) and a.is_integer():
            assert a == b
        else:
            assert abs(a - b) <= 2e-6 * abs(b) + 1e-10, (a, b)
    # header: the box to single precision
    ob = [float(x) for x in open(p).read().splitlines()[3].split()]
    rb = [float(x) for x in Z["restart"].tobytes().decode().splitlines()[3].split()]
    assert np.allclose(ob, rb, rtol=1e-7)


def test_restart_reads_back_as_a_coordinate_file(tmp_path):
    sysm = system.cg_melt(nchains=5, nbeads=20, seed=4, avoid=0.8)
    d = str(tmp_path / "case")
    system.write_dir(sysm, d)
    v_red = units.vel_to_reduced(sysm.v)
    engine.restart_write(os.path.join(d, "coordinate"), "again", 1.5, sysm.box, sysm.itype,
                         [len(j) for j in sysm.jbond], [(list(j) + [-1] * 4)[:4] for j in sysm.jbond], sysm.s,
                         v_red, sysm.mol_sizes, sysm.mol_names)
    back = system.read_dir(d)
    assert np.allclose(back.s, sysm.s, rtol=1e-13, atol=0)
    assert np.allclose(back.v, sysm.v, rtol=0, atol=6e-11)
    assert list(back.mol_sizes) == list(sysm.mol_sizes) and list(back.mol_names) == list(sysm.mol_names)
    assert [list(j) for j in back.jbond] == [list(j) for j in sysm.jbond]
    assert np.array_equal(back.itype, sysm.itype)
