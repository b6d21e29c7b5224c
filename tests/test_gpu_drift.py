"""NVE energy drift against the reference's own (the north-star gate).

The fixture cg10k_nve_drift holds an equilibrated 10,000-bead CG melt (200 x 50,
8 beads/nm^3, LJ 0.47 nm / 3.5 kJ/mol to 1.5 nm, list 1.8 nm every 10 steps,
5 fs, NVE) and the total energy of every step of 1,000 steps run by the
reference's own routines (MOVE.F90:40-88 in NEW_LOOP.F90:19-156 order, FP32).
The device path runs the same input for the same steps; the slope of a
straight-line fit to its total energy (V_NB + V_BOND + V_ANGLE + EK) per ps
must be no larger in magnitude than the reference's.
"""
import os
import sys

import numpy as np
import pytest

from ibisco_amd import system
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)


def test_nve_drift_no_larger_than_reference():
    from make_golden import drift_fit
    z = np.load(os.path.join(HERE, "cg10k_nve_drift.npz"))
    sy = system.from_arrays({k[4:]: z[k] for k in z.files if k.startswith("sys_")})
    n = int(z["nsteps"])
    ref = float(z["ref_drift_per_ps"])
    e = Engine(sy)
    etot = np.zeros(n)
    for k in range(n):
        e.step(1)
        ob = e.obs()
        etot[k] = sum(float(ob[q].sum()) for q in ("vnb", "vbond", "vangle", "vtors", "voop")) + float(ob["ek"].sum())
    dev = drift_fit(etot, sy.control.dt)
    print(f"NVE drift over {n} steps: device {dev:.4e}, reference {ref:.4e} per ps (reduced energy); "
          f"device E {etot[0]:.6e} -> {etot[-1]:.6e}, reference {z['ref_etot'][0]:.6e} -> {z['ref_etot'][-1]:.6e}",
          flush=True)
    assert np.all(np.isfinite(etot))
    assert abs(etot[0] - z["ref_etot"][0]) <= 1e-5 * abs(z["ref_etot"][0])   # the same start
    assert abs(dev) <= abs(ref)
