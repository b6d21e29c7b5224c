"""The equilibration protocol of the benchmarks, run through the FP64 oracle."""
import copy

import numpy as np

from ibisco_amd import system, units, workloads


class OracleEngine:
    """The subset of Engine that workloads.equilibrate drives, on the oracle."""

    def __init__(self, oracle_mod, sysm):
        self.mod = oracle_mod
        self.sysm = sysm
        self.table = (sysm.nb_tables[0].x, sysm.nb_tables[0].v)
        self.o = oracle_mod.Oracle(sysm)
        self.steps = 0

    def _rebuild(self):
        s, v, _ = self.o.state()
        sy = copy.copy(self.sysm)
        sy.control = copy.copy(self.sysm.control)
        sy.nb_tables = [system.Table((0, 0), self.table[0], self.table[1], "nb1")]
        self.o = self.mod.Oracle(sy)
        self.o.set_state(s[:sy.natoms], v)
        self.o.update_lists()

    def set_nb_table(self, x, v_reduced, sys=0, t=0):
        self.table = (np.asarray(x), units.reduced_to_kjmol(np.asarray(v_reduced)))
        self._rebuild()

    def set_thermostat(self, ensemble, temp_in, fac, dt):
        self.sysm = copy.copy(self.sysm)
        self.sysm.control = c = copy.copy(self.sysm.control)
        c.ensemble = "NVT" if ensemble == 1 else "NVE"
        c.temperature = temp_in * units.TEMPSCALE
        c.dt = dt * units.TIMESCALE / 1e-12
        c.taut = c.dt / fac
        self._rebuild()

    def step(self, n):
        self.o.run(n, self.steps)
        self.steps += n


def test_equilibrate_removes_overlaps(oracle_mod):
    s = system.ua_melt(nchains=10, nbeads=60, seed=2)
    target = s.control.temperature
    e = OracleEngine(oracle_mod, s)
    workloads.equilibrate(e, s, stage_steps=40, final_steps=80)
    ob = e.o.obs()
    pos, v, r = e.o.state()
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(v))
    assert 0.5 * target < ob["temp"] * units.TEMPSCALE < 2.0 * target
    # no pair closer than 0.6 sigma once relaxed
    d = r[:, None, :] - r[None, :, :]
    d -= np.round(d / s.box) * s.box
    dist = np.sqrt((d ** 2).sum(-1)) + np.eye(len(r)) * 10
    excluded = np.zeros_like(dist, dtype=bool)   # 1-2 and 1-3 partners do not interact (non_bonded 4)
    for i, js in enumerate(s.jbond):
        for j in js:
            excluded[i, j] = True
            for k in s.jbond[j]:
                excluded[i, k] = True
    assert dist[~excluded].min() > 0.6 * 0.401
