"""Iterative Boltzmann inversion over batched replicas.

IBIsCO exists to refine coarse-grained nonbonded potentials by iterative
Boltzmann inversion (citations.bib:27-40); the reference tree holds no code for
the outer loop, only the per-step force path it calls.  Here each iteration
runs every replica of one batched context with its own current nonbonded table,
reads each replica's radial distribution function (accumulated on the device by
the list build at every rebuild, so no per-step copies), and updates its table

    V_{n+1}(r) = V_n(r) + alpha * k_B T * ln(g_n(r) / g_target(r))

on the table's own abscissas; the new tables go back to the device with
ibx_set_nb_table, and the replicas continue from their current states.

The correction is formed on the RDF bins (centres (b + 1/2) dr), optionally
smoothed with [1, 2, 1]/4, taken as zero where the target has no weight
(g_target below ``gmin``, the excluded core: there the innermost valid
correction is carried inward), interpolated linearly onto the table rows, and
shifted so that it vanishes at the cutoff (the tables are shifted, V(rc) = 0).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import units

R_KJMOL = units.KB * units.NA / 1000.0   # k_B N_A in kJ/mol/K


def ibi_correction(r_tab, g, g_target, dr, temp_k, alpha=1.0, gmin=1e-2, smooth=1):
    """alpha * k_B T * ln(g / g_target) on the table abscissas r_tab (kJ/mol), zero at the last row."""
    g = np.asarray(g, np.float64)
    gt = np.asarray(g_target, np.float64)
    nb = min(len(g), len(gt))
    g, gt = g[:nb], gt[:nb]
    rb = (np.arange(nb) + 0.5) * dr
    ok = (gt > gmin) & (g > gmin)
    c = np.zeros(nb)
    c[ok] = alpha * R_KJMOL * temp_k * np.log(g[ok] / gt[ok])
    if not ok.any():
        return np.zeros_like(np.asarray(r_tab, np.float64))
    first = int(np.argmax(ok))
    c[:first] = c[first]                    # the core: the innermost valid correction carried inward
    for q in range(first + 1, nb):          # gaps beyond the core (g dips below gmin): the previous value
        if not ok[q]:
            c[q] = c[q - 1]
    for _ in range(max(0, smooth)):
        p = np.concatenate(([c[0]], c, [c[-1]]))
        c = 0.25 * p[:-2] + 0.5 * p[1:-1] + 0.25 * p[2:]
    r_tab = np.asarray(r_tab, np.float64)
    out = np.interp(r_tab, rb, c)           # constant beyond the first and last bin centres
    return out - out[-1]


def rdf_distance(g, g_target, gmin=1e-2):
    """Root-mean-square difference of two RDFs over the bins where the target has weight."""
    g = np.asarray(g, np.float64)
    gt = np.asarray(g_target, np.float64)
    nb = min(len(g), len(gt))
    m = gt[:nb] > gmin
    return float(np.sqrt(np.mean((g[:nb][m] - gt[:nb][m]) ** 2))) if m.any() else 0.0


@dataclass
class IbiResult:
    tables: list                      # per replica: (x, v) of the final raw nonbonded table (kJ/mol)
    distance: list = field(default_factory=list)   # per iteration: per replica RDF distance before the update
    rdf: list = field(default_factory=list)        # per replica: the last RDF


def run_ibi(engine, systems, g_targets, niter, nsteps, nequil=0, nbin=150, dr=0.01, alpha=0.5, gmin=1e-2,
            smooth=1, table=0, graph=True):
    """Iterative Boltzmann inversion of nonbonded table ``table`` of every replica of ``engine`` (whose systems
    are ``systems``, in order) towards ``g_targets[k]``: per iteration ``nequil`` steps after the table update,
    then ``nsteps`` steps sampling the RDF, then the update.  Runs at the temperature in each system's control
    (the caller sets the thermostat).  Returns the final tables and the distance history."""
    if len(g_targets) != len(systems):
        raise ValueError("one target RDF per replica")
    engine.rdf_enable(nbin, dr)
    x = [np.asarray(sy.nb_tables[table].x, np.float64) for sy in systems]
    v = [np.asarray(sy.nb_tables[table].v, np.float64).copy() for sy in systems]
    res = IbiResult(tables=[])
    for _ in range(niter):
        if nequil:
            engine.step(nequil, graph=graph)
        for k in range(len(systems)):
            engine.rdf(k, reset=True)       # discard what the equilibration steps added
        engine.step(nsteps, graph=graph)
        g = [engine.rdf(k, reset=True) for k in range(len(systems))]
        res.distance.append([rdf_distance(g[k], g_targets[k], gmin) for k in range(len(systems))])
        for k, sy in enumerate(systems):
            v[k] = v[k] + ibi_correction(x[k], g[k], g_targets[k], dr, sy.control.temperature, alpha, gmin,
                                         smooth)
            engine.set_nb_table(x[k], units.kjmol_to_reduced(v[k]), sys=k, t=table)
        res.rdf = g
    res.tables = list(zip(x, v))
    return res
