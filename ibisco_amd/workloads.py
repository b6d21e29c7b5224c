"""The benchmark workloads as synthetic systems, and their equilibration.

cg_melt_10k    CG melt, 200 chains x 50 beads (10,000), 8 beads/nm^3, 500 K, LJ(0.47, 3.5) to 1.5 nm
ua_melt_100k   UA melt, 400 x 250 CH2 (100,000), 0.78 g/cm^3, 450 K, LJ(0.401, 0.39) to 1.0 nm, RB torsion
cg_replicas    CG melts of 2,000-50,000 beads (log-uniform), chain length 20-100, each with its own
               Gaussian-corrected table
serving_jobs   a Poisson stream of such jobs (sizes 1,000-100,000, 500-5,000 steps) at 20 jobs/s

Chains are random walks (fixed bond length, bond angle near theta0) that may
overlap; :func:`equilibrate` removes the overlaps on the device with a capped
table under velocity rescaling, then relaxes with the real table under the
Berendsen thermostat of MOVE.F90:32-38.
"""
from __future__ import annotations

import math

import numpy as np

from . import system, units


def cg_melt_10k(seed=0, nchains=200, nbeads=50):
    return system.cg_melt(nchains=nchains, nbeads=nbeads, seed=seed, temp_k=500.0, dt=0.005, nupdate=10)


def ua_melt_100k(seed=1, nchains=400, nbeads=250, rlist=1.3, nupdate=10):
    return system.ua_melt(nchains=nchains, nbeads=nbeads, seed=seed, temp_k=450.0, dt=0.001, rlist=rlist,
                          nupdate=nupdate)


def cg_replicas(nrep=256, seed=3, rlist=1.8, nupdate=10, dt=0.005, scale=1.0):
    """nrep CG melt replicas of log-uniform size, each with its own table; ``scale`` shrinks every replica."""
    out = []
    for k, (nch, c) in enumerate(system.replica_sizes(nrep, seed)):
        nch = max(2, int(round(nch * scale)))
        r, v = system.corrected_cg_table(seed * 1000 + k)
        sy = system.cg_melt(nchains=nch, nbeads=c, seed=seed * 1000 + k, temp_k=500.0, dt=dt, rlist=rlist,
                            nupdate=nupdate, nb_table=(r, v))
        out.append(sy)
    return out


def capped_table(r, v, rcap):
    """The table with a straight-line continuation (constant force) below rcap."""
    v = np.asarray(v, dtype=np.float64).copy()
    dr = r[1] - r[0]
    kc = int(round(rcap / dr))
    if kc < 1 or kc >= len(r) - 1:
        return v
    slope = (v[kc + 1] - v[kc - 1]) / (2 * dr)
    v[:kc] = v[kc] + (r[:kc] - r[kc]) * slope
    return v


CAP_STAGES = (1.0, 0.95, 0.9, 0.85, 0.8, 0.75, 0.7, 0.65)


def equilibrate(engine, systems, stage_steps=100, final_steps=200, caps=CAP_STAGES):
    """Removes the overlaps of random-walk chains: the nonbonded table is capped
    (constant force below cap*sigma) at decreasing caps while velocities are
    rescaled to the target temperature every step (Berendsen with taut = dt),
    at half the production time step; then the real table relaxes under a
    weaker coupling.  Ends in the systems' own ensemble and time step."""
    if not isinstance(systems, (list, tuple)):
        systems = [systems]
    c0 = systems[0].control
    temp = units.temp_to_reduced(c0.temperature)
    dt = units.dt_to_reduced(c0.dt)
    engine.set_thermostat(1, temp, 1.0, 0.5 * dt)
    for cap in caps:
        for k, sy in enumerate(systems):
            for q, t in enumerate(sy.nb_tables):
                sig = _sigma_guess(t)
                engine.set_nb_table(t.x, units.kjmol_to_reduced(capped_table(t.x, t.v, cap * sig)), sys=k, t=q)
        engine.step(stage_steps)
    for k, sy in enumerate(systems):
        for q, t in enumerate(sy.nb_tables):
            engine.set_nb_table(t.x, units.kjmol_to_reduced(t.v), sys=k, t=q)
    engine.step(final_steps // 4)
    engine.set_thermostat(1, temp, 0.1, 0.5 * dt)
    engine.step(final_steps // 4)
    engine.set_thermostat(1, temp, dt / units.dt_to_reduced(0.1), dt)
    engine.step(final_steps - 2 * (final_steps // 4))
    ens = 0 if c0.ensemble.upper() == "NVE" else 1
    engine.set_thermostat(ens, temp, dt / units.dt_to_reduced(c0.taut), dt)


def _sigma_guess(t):
    """Where the tabulated potential first turns negative (LJ sigma)."""
    neg = np.nonzero(np.asarray(t.v) < 0)[0]
    return float(t.x[neg[0]]) if len(neg) else float(t.x[-1]) * 0.3


def serving_jobs(stream, cmin=20, cmax=100, rlist=1.8, nupdate=10, dt=0.005):
    """The systems of a serving stream: CG melts of about n_beads beads (chain
    length uniform in cmin..cmax), each with its own Gaussian-corrected table."""
    out = []
    for (_t, n, _steps, sd) in stream:
        rng = np.random.default_rng(sd)
        c = int(rng.integers(cmin, cmax + 1))
        nch = max(2, int(round(n / c)))
        r, v = system.corrected_cg_table(sd)
        out.append(system.cg_melt(nchains=nch, nbeads=c, seed=sd, temp_k=500.0, dt=dt, rlist=rlist,
                                  nupdate=nupdate, nb_table=(r, v)))
    return out


def serving_stream(rate_hz, duration_s, seed, nmin=1000, nmax=100000, smin=500, smax=5000):
    """Poisson arrivals of (t_arrive, n_beads, n_steps, table_seed)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    k = 0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            break
        n = int(math.exp(rng.uniform(math.log(nmin), math.log(nmax))))
        out.append((t, n, int(rng.integers(smin, smax + 1)), seed * 100000 + k))
        k += 1
    return out
