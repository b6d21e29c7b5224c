"""On-device RDF (no reference counterpart): histograms of the list pairs at each rebuild.

The list holds the nonbonded pairs, so the RDF leaves out excluded (bonded)
pairs, as an iterative Boltzmann inversion of the nonbonded potential wants.
Checked against a brute-force FP64 histogram of all non-excluded pairs at the
same positions (counts can differ only for a pair within FP32 rounding of a bin
edge), and a batch of two systems with different boxes against each alone.
"""
import numpy as np
import pytest

from ibisco_amd import system
from ibisco_amd.engine import Engine

pytestmark = pytest.mark.gpu

NBIN, DR = 150, 0.01


def _brute_g(s, box, excl_off, excl):
    n = len(s)
    d = s[:, None, :] - s[None, :, :]
    d -= np.round(d / box) * box
    r2 = (d * d).sum(-1)
    for i in range(n):
        r2[i, excl[excl_off[i]:excl_off[i + 1]]] = np.inf
    r = np.sqrt(r2[np.triu_indices(n, 1)])
    h = np.bincount((r[r < NBIN * DR] / DR).astype(int), minlength=NBIN)[:NBIN] * 2.0
    edges = np.arange(NBIN + 1) * DR
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    return h, h / (n * (n / np.prod(box)) * shell)


def test_rdf_matches_brute_force_pairs(oracle_mod):
    s = system.cg_melt(nchains=40, nbeads=50, seed=3, avoid=0.8, dt=0.002)
    excl_off, excl = oracle_mod.Oracle(s).exclusions()
    e = Engine(s)
    e.rdf_enable(NBIN, DR)
    e.step(9)
    s9, _ = e.state()
    e.step(1)   # step 10 rebuilds the list at the positions after step 9 and histograms it
    g = e.rdf(0)
    box = np.asarray(s.box)
    h_ref, g_ref = _brute_g(s9, box, excl_off, excl)
    n = len(s9)
    edges = np.arange(NBIN + 1) * DR
    h = g * n * (n / np.prod(box)) * 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    print("pairs", h_ref.sum() / 2, "max count difference", np.abs(h - h_ref).max())
    assert np.abs(h - h_ref).max() <= 2.0 + 1e-6
    assert abs(h.sum() - h_ref.sum()) <= 2.0 + 1e-6


def test_rdf_batch_of_different_boxes_equals_alone():
    a = system.cg_melt(nchains=24, nbeads=50, seed=11, avoid=0.8, dt=0.002)
    b = system.cg_melt(nchains=40, nbeads=50, seed=12, avoid=0.8, dt=0.002)
    assert a.box[0] != b.box[0]
    got = []
    for systems in ([a], [b], [a, b]):
        e = Engine(systems)
        e.rdf_enable(NBIN, DR)
        e.step(20, graph=True)
        got.append([e.rdf(k) for k in range(len(systems))])
        e.close()
    assert np.array_equal(got[0][0], got[2][0])
    assert np.array_equal(got[1][0], got[2][1])
    assert np.all(np.isfinite(got[2][1])) and got[2][1][-1] > 0
