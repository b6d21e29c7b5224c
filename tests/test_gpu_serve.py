"""Retirement of batch members and the serving loop (no reference counterpart).

A job's results from the serving loop (final state, observables, window sums,
mean pressure, RDF) must equal, bit for bit, those of the same system stepped
alone in its own context from step 0 with the RDF on, whichever batch it
shared, whenever it was admitted and however its batch was chunked.  Batching
is bit-exact because every per-system reduction and list run is independent of
the other systems; retirement only stops a system's workgroups.
"""
import os
import time

import numpy as np
import pytest

from ibisco_amd import system, units, workloads
from ibisco_amd.engine import Engine, IbxError, Server, estimate_bytes

pytestmark = pytest.mark.gpu

NBIN, DR = 150, 0.01
IBX_E_NOMEM, IBX_E_INVALID = -5, -1
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _job(nchains, nbeads, seed):
    r, v = system.corrected_cg_table(seed)
    return system.cg_melt(nchains=nchains, nbeads=nbeads, seed=seed, avoid=0.8, dt=0.002, nb_table=(r, v))


def _equilibrated(systems):
    """The systems' states after the capped-table relaxation, run as one batch."""
    e = Engine(systems)
    workloads.equilibrate(e, systems, stage_steps=50, final_steps=200)
    for k, sy in enumerate(systems):
        s, v = e.state(k)
        sy.s, sy.v = s[:sy.natoms].copy(), units.vel_from_reduced(v)
    e.close()
    return systems


def _alone(sysm, nsteps):
    e = Engine(sysm)
    e.rdf_enable(NBIN, DR)
    e.step(nsteps, graph=True)
    s, v = e.state()
    out = dict(s=s, v=v, rdf=e.rdf(0), sums=e.averages(0), obs=e.obs(0))
    e.close()
    return out


def test_retired_member_freezes_and_others_continue_bitwise():
    a, b = _equilibrated([_job(12, 40, 21), _job(20, 30, 22)])
    e = Engine([a, b])
    e.step(23, graph=True)
    e.retire(0)
    sa, va = e.state(0)
    oa = e.obs(0)
    e.step(40)   # steps 24..63 eager: a is frozen, b goes on
    sa2, va2 = e.state(0)
    sb, vb = e.state(1)
    oa2 = e.obs(0)
    e.close()
    assert np.array_equal(sa, sa2) and np.array_equal(va, va2)
    assert all(np.array_equal(oa[k], oa2[k]) for k in ("vnb", "pt", "ek"))
    ea, eb = Engine(a), Engine(b)
    ea.step(23, graph=True)
    eb.step(63, graph=True)
    ra, rb = ea.state(), eb.state()
    ea.close()
    eb.close()
    assert np.all(np.isfinite(sb))
    assert np.array_equal(sa, ra[0]) and np.array_equal(va, ra[1])
    assert np.array_equal(sb, rb[0]) and np.array_equal(vb, rb[1])


@pytest.mark.parametrize("case", ["cg", "ua", "hybridPA"])
def test_footprint_estimate_matches_allocation(case):
    if case == "cg":
        sysm = _job(20, 50, 5)
    elif case == "ua":
        sysm = system.ua_melt(nchains=16, nbeads=60, seed=2)
    else:
        z = np.load(os.path.join(HERE, "hybridPA.npz"))
        sysm = system.from_arrays({k[4:]: z[k] for k in z.files if k.startswith("sys_")})
        sysm.control.ensemble = "NVE"
    est = estimate_bytes(sysm, NBIN)
    e = Engine(sysm)
    e.rdf_enable(NBIN, DR)
    got = e.info(0)[5]
    e.close()
    print(case, "estimate", est, "allocated", got, flush=True)
    assert got <= est <= got * 1.01


def test_served_jobs_equal_alone():
    specs = [(10, 40, 31, 180), (14, 30, 32, 23), (8, 50, 33, 57), (16, 25, 34, 40), (12, 35, 35, 75),
             (6, 60, 36, 31), (9, 44, 37, 66)]
    systems = _equilibrated([_job(nc, nb, sd) for nc, nb, sd, _ in specs])
    srv = Server(max_cohorts=2, max_cohort_jobs=3, chunk_steps=16, rdf_nbin=NBIN, rdf_dr=DR, step_band=0.0)
    ids = []
    for k, (sy, sp) in enumerate(zip(systems, specs)):
        ids.append(srv.submit(sy, sp[3], keep_state=True))
        if k == 2:
            time.sleep(0.05)   # later jobs arrive while the first batches run
    res = [srv.wait(j, timeout=120) for j in ids]
    st = srv.stats()
    srv.close()
    print("stats", st, flush=True)
    assert all(r is not None for r in res)
    assert st["completed"] == len(specs) and st["failed"] == 0
    assert max(r["cohort_jobs"] for r in res) >= 2     # some jobs shared a batch
    for sy, sp, r in zip(systems, specs, res):
        ref = _alone(sy, sp[3])
        print("job", sp, "cohort", r["cohort"], "of", r["cohort_jobs"], flush=True)
        assert r["nsteps"] == sp[3] and r["sums"][28] == sp[3]
        assert np.array_equal(r["s"], ref["s"]) and np.array_equal(r["v"], ref["v"])
        assert np.array_equal(r["rdf"], ref["rdf"])
        assert np.array_equal(r["sums"][:29], ref["sums"][:29])
        p = (ref["sums"][15] + ref["sums"][16] + ref["sums"][17]) / (3 * ref["sums"][28])
        assert r["mean_pressure"] == p
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]


def test_step_band_groups_similar_jobs():
    """With a step band, a new batch holds the oldest queued job and only jobs whose step counts are within
    the band of its own (here with hand-overs of free places off, so that batches take jobs only when they
    are created); every result still equals the job alone."""
    specs = [(8, 40, 41, 200), (8, 40, 42, 30), (8, 40, 43, 190), (8, 40, 44, 35), (8, 40, 45, 210)]
    *systems, slow = _equilibrated([_job(nc, nb, sd) for nc, nb, sd, _ in specs] + [_job(4, 40, 49)])
    srv = Server(max_cohorts=1, max_cohort_jobs=8, chunk_steps=64, rdf_nbin=NBIN, rdf_dr=DR, step_band=1.5,
                 spare_places=-1)
    blocker = srv.submit(slow, 400)     # keeps the only batch slot busy while the rest queue up
    ids = [srv.submit(sy, sp[3], keep_state=True) for sy, sp in zip(systems, specs)]
    res = [srv.wait(j, timeout=120) for j in ids]
    srv.wait(blocker, timeout=120)
    srv.close()
    cohort = {sp[3]: r["cohort"] for sp, r in zip(specs, res)}
    print("cohorts by step count", cohort, flush=True)
    assert cohort[200] == cohort[190] == cohort[210]
    assert cohort[30] == cohort[35] != cohort[200]
    for sy, sp, r in zip(systems, specs, res):
        ref = _alone(sy, sp[3])
        assert np.array_equal(r["s"], ref["s"]) and np.array_equal(r["rdf"], ref["rdf"])


def test_last_step_stops_a_member_on_the_device():
    """A member given a last step (ibx_set_last_step) stops by itself there while the batch steps on in whole
    graph blocks: its state, observables, window sums and RDF equal the system alone after that many steps,
    and the other member equals itself alone after all of them."""
    a, b = _equilibrated([_job(12, 40, 81), _job(10, 30, 82)])
    e = Engine([a, b])
    e.rdf_enable(NBIN, DR)
    e.set_last_step(1, 27)
    e.step(40, graph=True)
    e.step(10, graph=True)
    sb, vb = e.state(1)
    rb = dict(rdf=e.rdf(1), sums=e.averages(1), obs=e.obs(1))
    sa, va = e.state(0)
    cnt = e.counters()
    e.close()
    assert cnt["eager_steps"] == 0
    ref = _alone(b, 27)
    assert np.array_equal(sb, ref["s"]) and np.array_equal(vb, ref["v"])
    assert np.array_equal(rb["rdf"], ref["rdf"]) and np.array_equal(rb["sums"][:29], ref["sums"][:29])
    assert rb["sums"][28] == 27 and all(np.array_equal(rb["obs"][k], ref["obs"][k]) for k in ("vnb", "pt", "ek"))
    ref = _alone(a, 50)
    assert np.array_equal(sa, ref["s"]) and np.array_equal(va, ref["v"])


def _running(srv, timeout=60.0):
    t = time.time()
    while srv.stats()["running"] < 1:
        assert time.time() - t < timeout
        time.sleep(0.01)


def test_jobs_join_a_running_batch():
    """Continuous batching: jobs that arrive while the only batch runs are given its spare places, and then
    the place a finished job left, each at a block boundary; every result equals the job alone, bit for bit."""
    long, *short = _equilibrated([_job(50, 40, 61), _job(8, 30, 62), _job(10, 25, 63), _job(9, 30, 64)])
    nlong = 40000
    srv = Server(max_cohorts=1, chunk_steps=50, rdf_nbin=NBIN, rdf_dr=DR, step_band=0.0, spare_places=2)
    jl = srv.submit(long, nlong, keep_state=True)
    _running(srv)
    time.sleep(0.2)
    steps = (37, 52, 29)
    ids = [srv.submit(sy, n, keep_state=True) for sy, n in zip(short, steps)]
    res = [srv.wait(j, timeout=120) for j in ids]
    rl = srv.wait(jl, timeout=120)
    st = srv.stats()
    srv.close()
    print("stats", st, flush=True)
    print("long job", rl["t_admit"], rl["t_done"], [(r["t_admit"], r["t_done"]) for r in res], flush=True)
    assert st["cohorts"] == 1 and st["admitted_running"] == 3 and st["failed"] == 0
    assert all(r["cohort"] == rl["cohort"] and r["t_admit"] < rl["t_done"] for r in res)
    # two spare places for three newcomers: the third started after one of the first two had finished
    assert res[2]["t_admit"] >= min(res[0]["t_done"], res[1]["t_done"])
    for sy, n, r in zip(short + [long], steps + (nlong,), res + [rl]):
        ref = _alone(sy, n)
        assert np.array_equal(r["s"], ref["s"]) and np.array_equal(r["v"], ref["v"])
        assert np.array_equal(r["rdf"], ref["rdf"]) and np.array_equal(r["sums"][:29], ref["sums"][:29])


def test_job_fitting_no_place_gets_a_new_batch():
    """A job larger than every free place, with no room for a second batch, waits until the running batch,
    which then takes no more jobs, has ended; it then runs in a batch of its own and equals the job alone."""
    small, big, late = _equilibrated([_job(8, 40, 71), _job(60, 50, 72), _job(8, 30, 73)])
    srv = Server(max_cohorts=1, chunk_steps=50, rdf_nbin=NBIN, rdf_dr=DR, step_band=0.0, spare_places=1)
    js = srv.submit(small, 100000)
    _running(srv)
    jb = srv.submit(big, 60, keep_state=True)       # 3,000 beads: no place of the running batch holds it
    time.sleep(0.5)
    jl = srv.submit(late, 40, keep_state=True)      # fits the spare place, but the batch is draining
    rs, rb, rl = srv.wait(js, timeout=120), srv.wait(jb, timeout=120), srv.wait(jl, timeout=120)
    st = srv.stats()
    srv.close()
    print("stats", st, [(r["cohort"], r["t_admit"], r["t_done"]) for r in (rs, rb, rl)], flush=True)
    assert st["failed"] == 0 and st["cohorts"] == 2 and st["admitted_running"] == 0
    assert rb["cohort"] != rs["cohort"] and rb["t_admit"] >= rs["t_done"]
    for sy, n, r in ((big, 60, rb), (late, 40, rl)):
        ref = _alone(sy, n)
        assert np.array_equal(r["s"], ref["s"]) and np.array_equal(r["rdf"], ref["rdf"])


def test_assigned_place_equals_alone():
    """A place of a running batch, retired, is given to another system at a list-interval boundary
    (ibx_assign): the newcomer's state, observables, window sums and RDF equal the system stepped alone from
    step 0, bit for bit, and so do those of the system that ran on beside it; a second newcomer reuses the
    place again; the graphs captured before the hand-overs are replayed after them, with no new capture."""
    a, b, c, d = _equilibrated([_job(12, 40, 51), _job(20, 30, 52), _job(10, 40, 53), _job(9, 30, 54)])
    e = Engine([a, b])
    e.rdf_enable(NBIN, DR)
    cap = e.capacity(1)
    print("place 1 holds", cap, flush=True)
    assert cap["period"] == 10
    e.step(40, graph=True)                         # captures the 40-step block
    e.step(10, graph=True)                         # and the 10-step block
    e.step(3, graph=True)
    e.retire(1)                                    # b's job ends at its step 53
    e.step(7, graph=True)                          # to step 60, a multiple of the block period
    with pytest.raises(IbxError) as ei:            # a live place is not given away
        e.assign(0, c)
    assert ei.value.code == IBX_E_INVALID
    with pytest.raises(IbxError) as ei:            # a newcomer starts on a list-interval boundary
        e.assign(1, c, step0=e.step_count + 1)
    assert ei.value.code == IBX_E_INVALID
    with pytest.raises(IbxError) as ei:            # a system larger than the place
        e.assign(1, _job(30, 40, 55))
    assert ei.value.code == IBX_E_NOMEM
    captures = e.counters()["captures"]
    e.assign(1, c)
    e.step(47, graph=True)                         # c: its steps 1..47 (a 40-step block, 7 eager steps)
    sc, vc = e.state(1)
    rc = dict(rdf=e.rdf(1), sums=e.averages(1), obs=e.obs(1))
    e.retire(1)
    e.step(3, graph=True)                          # to step 110
    e.assign(1, d)
    e.step(35, graph=True)                         # d: its steps 1..35 (three 10-step blocks); a: 145 in all
    sd, vd = e.state(1)
    rd = dict(rdf=e.rdf(1), sums=e.averages(1), obs=e.obs(1))
    sa, va = e.state(0)
    ra = dict(rdf=e.rdf(0), sums=e.averages(0))
    cnt = e.counters()
    e.close()
    print("counters", cnt, "captures before the first hand-over", captures, flush=True)
    assert captures == 2 and cnt["captures"] == captures and cnt["graph_steps"] == 40 + 10 + 40 + 30
    for got_s, got_v, got, sy, n in ((sc, vc, rc, c, 47), (sd, vd, rd, d, 35)):
        ref = _alone(sy, n)
        assert np.array_equal(got_s, ref["s"]) and np.array_equal(got_v, ref["v"])
        assert np.array_equal(got["rdf"], ref["rdf"])
        assert np.array_equal(got["sums"][:29], ref["sums"][:29]) and got["sums"][28] == n
        assert all(np.array_equal(got["obs"][k], ref["obs"][k]) for k in ("vnb", "pt", "ek"))
    ref = _alone(a, 145)
    assert np.array_equal(sa, ref["s"]) and np.array_equal(va, ref["v"])
    assert np.array_equal(ra["rdf"], ref["rdf"]) and np.array_equal(ra["sums"][:29], ref["sums"][:29])
