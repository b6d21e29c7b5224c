"""Builds libibx.so (HIP kernels for gfx950 + host runtime + C ABI) in-tree.

Usage: python -m ibisco_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libibx.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["kernels.hip", "api.hip", "topology.cpp", "io.cpp", "serve.cpp", "output.cpp"]
ARCH = os.environ.get("IBX_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I" + os.path.join(ROOT, "include"),
         "-I" + CSRC, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
         "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "c++"] + [f for f in FLAGS if not f.startswith("--offload-arch")
                                      and f not in ("-Xclang", "-target-feature", "-packed-fp32-ops")] + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(HERE, "ibx_driver")


def build_driver(verbose=False):
    """The Fortran step driver (iso_c_binding over libibx.so), when amdflang is installed."""
    fc = os.environ.get("FC", "amdflang")
    if subprocess.run(["which", fc], capture_output=True).returncode != 0:
        if verbose:
            print("no Fortran compiler: ibx_driver not built")
        return None
    mod = os.path.join(OBJDIR, "ibx_mod.o")
    cmds = [[fc, "-O2", "-c", os.path.join(FORTRAN, "ibx_mod.f90"), "-module-dir", OBJDIR, "-o", mod],
            [fc, "-O2", "-I" + OBJDIR, os.path.join(FORTRAN, "ibx_driver.f90"), mod, "-L" + HERE, "-libx",
             "-Wl,-rpath,$ORIGIN", "-o", DRIVER]]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJDIR)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_driver(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
