"""Iterative Boltzmann inversion over batched replicas on the GPU (no reference counterpart: the reference has
no code for the outer loop; its purpose is stated in citations.bib:27-40).

Three CG melt replicas, each with its own target: the RDF of the same melt under its own Gaussian-corrected table.
Starting from the plain LJ table, five table updates (V += 0.5 k_B T ln(g / g_target)) must bring every
replica's RDF much closer to its target, and the replicas' tables must move apart.
"""
import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine
from ibisco_amd.ibi import rdf_distance, run_ibi

pytestmark = pytest.mark.gpu

NBIN, DR = 150, 0.01
SEEDS = [31, 32, 33]


def _melt(seed, table):
    return system.cg_melt(nchains=60, nbeads=50, seed=seed, temp_k=500.0, dt=0.005, rlist=1.8, nupdate=10,
                          nb_table=table)


def _nvt(e, sy):
    dt = units.dt_to_reduced(sy.control.dt)
    e.set_thermostat(1, units.temp_to_reduced(sy.control.temperature), dt / units.dt_to_reduced(0.1), dt)


def test_ibi_brings_each_replica_to_its_target():
    ref = [_melt(s, system.corrected_cg_table(s)) for s in SEEDS]
    e = Engine(ref)
    workloads.equilibrate(e, ref, stage_steps=100, final_steps=400)
    _nvt(e, ref[0])
    e.rdf_enable(NBIN, DR)
    e.step(1000, graph=True)
    for k in range(len(ref)):
        e.rdf(k, reset=True)
    e.step(4000, graph=True)
    targets = [e.rdf(k) for k in range(len(ref))]
    states = [e.state(k) for k in range(len(ref))]
    e.close()

    lj = system.lj_sf_table(0.47, 3.5, 1.5, 0.002)
    start = [_melt(s, lj) for s in SEEDS]
    for sy, (s, v) in zip(start, states):
        sy.s, sy.v = s[:sy.natoms].copy(), units.vel_from_reduced(v)
    e = Engine(start)
    _nvt(e, start[0])
    res = run_ibi(e, start, targets, niter=6, nsteps=3000, nequil=500, alpha=0.5)
    e.close()
    d = np.array(res.distance)
    print("RDF distance per iteration and replica:\n", d, flush=True)
    assert np.all(np.isfinite(d))
    assert np.all(d[-1] < 0.5 * d[0]), d
    v = [t[1] for t in res.tables]
    assert not np.allclose(v[0], v[1]) and not np.allclose(v[1], v[2])
    assert all(rdf_distance(res.rdf[k], targets[k]) == d[-1][k] for k in range(len(SEEDS)))
