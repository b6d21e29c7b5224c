"""The kernels' unconditional loads stay in bounds on edge-case layouts.

tests/index_model.cpp replays, on the host, the loads the list build, the
integrator and the bonded terms issue before testing whether they need them,
with the index functions the kernels use (ibisco_amd/csrc/ibx_index.h) and
arrays of the sizes the library allocates; it is built with AddressSanitizer,
so an index that leaves its array fails the run.  A planted mistake (the
exclusion fallback left unclamped) must fail it too.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path_factory.mktemp("index_model") / "index_model")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "ibisco_amd", "csrc"), os.path.join(HERE, "index_model.cpp"), "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_unconditional_loads_in_bounds(model):
    r = subprocess.run([model], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and "index model ok" in r.stdout


def test_planted_unclamped_index_is_caught(model):
    r = subprocess.run([model, "--plant"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr
