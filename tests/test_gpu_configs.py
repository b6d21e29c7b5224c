"""The benchmark workloads against the FP64 oracle on the GPU.

- Batched replicas (the replicas workload): eight CG melts of 500-50,000 beads,
  each with its own Gaussian-corrected table (the smallest has fewer than 7
  list cells along x and y, so its members scan the whole system, and fewer than 3
  reference cells, so the oracle uses its all-pairs build), equilibrated
  together on the device, then each replica of the batch against the oracle on
  the same positions: pair sets bit-exact, forces <= 1e-5 relative RMS,
  energies and virial <= 1e-7; and 20 batched steps of three replicas against
  the oracle's own run (NONBONDED_FORCE.F90:45-104, NEW_LOOP.F90:19-156).
- Served jobs (the serving workload): jobs submitted to a serving loop of two
  batches in two waves, with jobs retiring at different steps; each job's
  final state, mean pressure (AVERAGE.f90:15) and RDF against an oracle run of
  that job alone (the RDF from the oracle's own lists at every rebuild).
- Full size: the 10,000-bead CG melt and the 100,000-site UA melt, pairs,
  forces, energies and virial against the oracle.
"""
import copy

import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine, Server

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-7
TERMS = ("vnb", "vbond", "vangle", "vtors", "voop")

# (chains, beads per chain): 500 ... 50,000 beads
REPLICAS = [(10, 50), (100, 20), (49, 100), (150, 60), (200, 80), (300, 90), (500, 100), (100, 30)]


def _replicas():
    out = []
    for k, (nch, c) in enumerate(REPLICAS):
        r, v = system.corrected_cg_table(700 + k)
        out.append(system.cg_melt(nchains=nch, nbeads=c, seed=700 + k, temp_k=500.0, dt=0.005, rlist=1.8,
                                  nupdate=10, nb_table=(r, v)))
    return out


def _with_state(sy, s, v):
    one = copy.copy(sy)
    one.s, one.v = s[:sy.natoms].copy(), units.vel_from_reduced(v)
    return one


def _check_against_oracle(oracle_mod, e, k, sy, label):
    """Replica k of context e (forces just computed with a fresh list) against the oracle on its positions."""
    s, v = e.state(k)
    o = oracle_mod.Oracle(_with_state(sy, s, v))
    o.update_lists()
    o.new_force()
    po, pg = o.pairs(0), e.pairs(k, 0)
    assert po.shape == pg.shape and np.array_equal(po, pg), (label, po.shape, pg.shape)
    fo, fg = o.forces(), e.forces(k)
    rr = float(np.linalg.norm(fg - fo) / np.linalg.norm(fo))
    oo, og = o.obs(), e.obs(k)
    eo = np.concatenate([oo[t] for t in TERMS])
    eg = np.concatenate([og[t] for t in TERMS])
    scale = np.abs(eo).sum()
    vir = np.abs(og["pt"] - oo["pt"]).max() / abs(oo["pt"][:3].sum())
    print(f"{label}: {sy.natoms} beads, {len(po)} pairs, force rel RMS {rr:.2e}, "
          f"energy {np.abs(eg - eo).max() / scale:.2e}, virial {vir:.2e}", flush=True)
    assert rr <= FORCE_RTOL, (label, rr)
    assert np.all(np.abs(eg - eo) <= ENERGY_RTOL * scale), (label, eg, eo)
    assert vir <= ENERGY_RTOL, (label, og["pt"], oo["pt"])


def test_batched_replicas_against_oracle(oracle_mod):
    systems = _replicas()
    e = Engine(systems)
    workloads.equilibrate(e, systems, stage_steps=100, final_steps=200)
    info = e.info(0)
    print("replicas", [s.natoms for s in systems], "padded items", info[7], flush=True)
    e.compute_forces(rebuild=True)
    for k, sy in enumerate(systems):
        _check_against_oracle(oracle_mod, e, k, sy, f"replica {k}")
    # 20 batched steps of the three smallest replicas against the oracle's run of each
    small = sorted(range(len(systems)), key=lambda k: systems[k].natoms)[:3]
    starts = {k: e.state(k) for k in small}
    step0 = e.step_count
    e.step(20, graph=True)
    for k in small:
        o = oracle_mod.Oracle(_with_state(systems[k], *starts[k]))
        o.run(20, step0)
        so, vo, _ = o.state()
        sg, vg = e.state(k)
        n = systems[k].natoms
        ds = np.abs(sg[:n] - so[:n]).max()
        dv = float(np.linalg.norm(vg - vo) / np.linalg.norm(vo))
        print(f"replica {k}: 20 steps, max |ds| {ds:.2e} nm, velocity rel {dv:.2e}", flush=True)
        assert ds <= 1e-6 and dv <= 1e-5


def _oracle_job(oracle_mod, sy, nsteps, nbin, dr):
    """An oracle run of one job from step 0: final state, mean of (PT11+PT22+PT33)/3 over its steps, and the
    RDF of the group-0 list pairs at every rebuild (normalised as the device does)."""
    o = oracle_mod.Oracle(sy)
    L = sy.box
    hist = np.zeros(nbin)
    nsamp = 0
    ptr = []
    nupdate = sy.control.nupdate
    for step in range(1, nsteps + 1):
        if step % nupdate == 0:   # the list this step builds, from the positions it starts with
            _, _, r = o.state()
            o.set_state()         # folds the positions as the step's SHIFT does
            o.update_lists()
            _, _, r = o.state()
            p = o.pairs(0)
            d = r[p[:, 0]] - r[p[:, 1]]
            d -= np.round(d / L) * L
            rr = np.sqrt((d * d).sum(1))
            b = np.floor(rr / dr).astype(np.int64)
            b = b[b < nbin]
            hist += np.bincount(b, minlength=nbin)[:nbin] * 2.0   # both members' rows
            nsamp += 1
        o.run(1, step - 1)
        ptr.append(o.obs()["pt"][:3].sum() / 3.0)
    n = sy.natoms
    vol = L[0] * L[1] * L[2]
    r0 = np.arange(nbin) * dr
    shell = 4.0 / 3.0 * np.pi * ((r0 + dr) ** 3 - r0 ** 3)
    g = hist / (max(nsamp, 1) * n * (n / vol) * shell)
    s, v, _ = o.state()
    return s, v, float(np.mean(ptr)), g


def test_served_jobs_against_oracle(oracle_mod):
    """Seven CG jobs (500-3,000 beads, 23-61 steps, own tables) in two waves to a serving loop of at most two
    batches: jobs retire at different steps while others continue; each result against the oracle."""
    specs = [(10, 50, 23), (30, 40, 61), (25, 60, 37), (12, 100, 44), (40, 50, 29), (20, 30, 53), (15, 70, 41)]
    jobs = []
    for k, (nch, c, n) in enumerate(specs):
        r, v = system.corrected_cg_table(900 + k)
        sy = system.cg_melt(nchains=nch, nbeads=c, seed=900 + k, temp_k=500.0, dt=0.005, rlist=1.8, nupdate=10,
                            nb_table=(r, v))
        jobs.append((sy, n))
    # relaxed starting states (a raw random-walk melt blows up within a few tens of steps)
    pre = Engine([sy for sy, _ in jobs])
    workloads.equilibrate(pre, [sy for sy, _ in jobs], stage_steps=100, final_steps=200)
    jobs = [(_with_state(sy, *pre.state(k)), n) for k, (sy, n) in enumerate(jobs)]
    pre.close()
    nbin, dr = 150, 0.01
    srv = Server(max_cohorts=2, chunk_steps=20, rdf_nbin=nbin, rdf_dr=dr, step_band=0.0)
    ids = [srv.submit(sy, n, keep_state=True) for sy, n in jobs[:4]]
    first = srv.wait(ids[0])
    ids += [srv.submit(sy, n, keep_state=True) for sy, n in jobs[4:]]
    res = [first] + [srv.wait(j) for j in ids[1:]]
    st = srv.stats()
    srv.close()
    print("server", st, flush=True)
    assert st["failed"] == 0 and st["cohorts"] >= 2
    for k, ((sy, n), r) in enumerate(zip(jobs, res)):
        so, vo, po, go = _oracle_job(oracle_mod, sy, n, nbin, dr)
        na = sy.natoms
        ds = np.abs(r["s"][:na] - so[:na]).max()
        dv = float(np.linalg.norm(r["v"] - vo) / np.linalg.norm(vo))
        dp = abs(r["mean_pressure"] - po) / abs(po)
        dg = np.abs(r["rdf"] - go).max()
        print(f"job {k}: {na} beads, {n} steps, cohort {r['cohort']} of {r['cohort_jobs']}: |ds| {ds:.2e}, "
              f"v {dv:.2e}, mean pressure {dp:.2e}, RDF max diff {dg:.2e} (max g {go.max():.2f})", flush=True)
        assert ds <= 1e-5 and dv <= 1e-4
        assert dp <= 1e-5
        assert dg <= 5e-3 * max(1.0, go.max())


@pytest.mark.parametrize("which", ["cg_melt_10k", "ua_melt_100k"])
def test_full_size_against_oracle(oracle_mod, which):
    sy = getattr(workloads, which)()
    e = Engine(sy)
    workloads.equilibrate(e, sy, stage_steps=200, final_steps=1000)   # removes the random walks' overlaps
    e.compute_forces(rebuild=True)
    _check_against_oracle(oracle_mod, e, 0, sy, which)
