"""Input errors of ibx_load_dir (RDCONTROL.f90:525-536, RDCOOR, RDINTERACT): a missing
control keyword, a missing file or a malformed table are reported as error
codes with the reference's message before any device work, so these run on the
CPU.  The device-side error paths are in test_gpu_errors.py."""
import os

import pytest

from ibisco_amd import engine, system

IBX_E_INVALID, IBX_E_IO = -1, -6


def _dir_with(tmp_path, edit):
    d = str(tmp_path)
    system.write_dir(system.cg_melt(nchains=4, nbeads=20, seed=1), d)
    edit(d)
    return d


def _load(d):
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libibx.so not built")
    with pytest.raises(engine.IbxError) as ei:
        engine.Engine(None, load_dir=d)
    return ei.value


def _drop_line(d, fname, key):
    p = os.path.join(d, fname)
    lines = [ln for ln in open(p) if not ln.split() or ln.split()[0] != key]
    open(p, "w").writelines(lines)


@pytest.mark.parametrize("key", ["cutoff", "time_step", "update_neighbour_list"])
def test_missing_control_keyword(tmp_path, key):
    e = _load(_dir_with(tmp_path, lambda d: _drop_line(d, "control", key)))
    assert e.code == IBX_E_INVALID and f"missing control keyword {key}" in str(e)


@pytest.mark.parametrize("fname", ["control", "interaction", "coordinate"])
def test_missing_file(tmp_path, fname):
    e = _load(_dir_with(tmp_path, lambda d: os.remove(os.path.join(d, fname))))
    assert e.code == IBX_E_IO and "does not exist" in str(e)


def test_invalid_ensemble(tmp_path):
    def edit(d):
        p = os.path.join(d, "control")
        txt = open(p).read().replace("NVE", "LA")
        open(p, "w").write(txt)
    e = _load(_dir_with(tmp_path, edit))
    assert e.code == IBX_E_INVALID and "Ensemble" in str(e)


def test_missing_table_file(tmp_path):
    def edit(d):
        for f in os.listdir(d):
            if f.startswith("nb"):
                os.remove(os.path.join(d, f))
    e = _load(_dir_with(tmp_path, edit))
    assert e.code == IBX_E_IO and "does not exist" in str(e)


@pytest.mark.parametrize("key", ["pressure", "pressure_coupling_time", "isothermal_compressibility"])
def test_npt_needs_barostat_keywords(tmp_path, key):
    """RDCONTROL.f90:307-345: NPT requires the barostat's keywords (NVE/NVT do not)."""
    def edit(d):
        p = os.path.join(d, "control")
        lines = []
        for ln in open(p):
            tok = ln.split()
            if tok and tok[0] == "ensemble":
                ln = "ensemble NPT\n"
            if tok and tok[0] == key:
                continue
            lines.append(ln)
        open(p, "w").writelines(lines)
    e = _load(_dir_with(tmp_path, edit))
    assert e.code == IBX_E_INVALID and f"missing control keyword {key}" in str(e)
