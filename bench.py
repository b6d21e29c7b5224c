"""Throughput of the per-step force + integration path on one MI355X.

Default workload: the united-atom melt (400 chains x 250 CH2 = 100,000 sites,
LJ 0.401 nm / 0.39 kJ/mol to 1.0 nm, harmonic bonds and angles, Ryckaert-
Bellemans torsion table, 450 K, 1 fs, list 1.3 nm rebuilt every 10 steps),
synthetic and equilibrated on the device before timing.  One "step" is one
full NEW_LOOP step (list rebuild when due, nonbonded, bonded, leap-frog,
observables) advanced through HIP graphs.

Prints one JSON line: whole-job atom-steps/s, the nonbonded kernel's
algorithmic-byte roofline fraction (HIP events on the library's stream), and
the reference CPU path (oracle/_ref, the reference's own routines built from
source) timed on this host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import resource
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s on 1 GPU, batched/served vs plain, at ≤1e-5 rel RMS force err"
HBM_PEAK_GBS = 8000.0


def _stack_unlimited():
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    except (ValueError, OSError):
        pass


def cpu_baseline(sysm, nsteps, threads):
    """Times the reference's own routines (oracle/_ref/ref_harness) on this host."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    from ibisco_amd import system
    d = tempfile.mkdtemp(prefix="ibx_ref_")
    try:
        system.write_dir(sysm, d)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
                   OMP_STACKSIZE="256M")
        t0 = time.time()
        r = subprocess.run([exe, "time", str(nsteps)], cwd=d, env=env, capture_output=True, text=True,
                           preexec_fn=_stack_unlimited, timeout=600)
        wall = time.time() - t0
        line = [x for x in r.stdout.splitlines() if x.startswith("LOOP_SECONDS")]
        if r.returncode != 0 or not line:
            return {"error": f"reference harness rc={r.returncode}: {r.stderr[-300:]}"}
        secs = float(line[0].split()[1])
        return {"value": sysm.natoms * nsteps / secs, "unit": "atom-steps/s", "cores": threads, "kind": "reference",
                "sample": f"{nsteps} steps of the same {sysm.natoms}-site system (list rebuilt at step 10), "
                          f"loop {secs:.2f} s, wall incl. set-up {wall:.1f} s, OpenMP FP32 reference routines"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def replicas_compare(nrep, steps, warmup, seed=3):
    """CG melt replicas (log-uniform 2k-50k beads, each with its own corrected table): atom-steps/s of all
    replicas in one context (every kernel launch covers all of them, HIP graphs of four list intervals) against
    the plain path, each replica alone in its own context, one after the other, from the same equilibrated
    states.  RDFs accumulate on the device at every list rebuild in the batched run."""
    import copy
    from ibisco_amd import units, workloads
    from ibisco_amd.engine import Engine

    systems = workloads.cg_replicas(nrep=nrep, seed=seed)
    beads = [s.natoms for s in systems]
    total = sum(beads)
    print(f"[bench] {nrep} replicas, {total} beads (min {min(beads)}, max {max(beads)})", file=sys.stderr, flush=True)
    eb = Engine(systems)
    workloads.equilibrate(eb, systems, stage_steps=100, final_steps=400)
    print("[bench] replicas equilibrated", file=sys.stderr, flush=True)
    eb.rdf_enable(150, 0.01, every=1)
    eb.step(warmup, graph=True)
    eb.obs()
    t0 = time.perf_counter()
    eb.step(steps, graph=True)
    eb.obs()
    tb = time.perf_counter() - t0
    states = [eb.state(k) for k in range(nrep)]
    eb.close()
    tp = 0.0
    for k, sy in enumerate(systems):
        one = copy.copy(sy)
        one.s, one.v = states[k][0][:sy.natoms].copy(), units.vel_from_reduced(states[k][1])
        e = Engine(one)
        e.step(warmup, graph=True)
        e.obs()
        t0 = time.perf_counter()
        e.step(steps, graph=True)
        e.obs()
        tp += time.perf_counter() - t0
        e.close()
        if (k + 1) % 32 == 0:
            print(f"[bench] plain path: {k + 1}/{nrep} replicas", file=sys.stderr, flush=True)
    vb, vp = total * steps / tb, total * steps / tp
    print(f"[bench] replicas: batched {vb:.3e}, plain {vp:.3e} atom-steps/s", file=sys.stderr, flush=True)
    return {"workload": "cg_replicas_corrected_tables", "replicas": nrep, "beads": total, "steps": steps,
            "batched": {"value": vb, "ms_per_step": 1e3 * tb / steps},
            "plain": {"value": vp, "ms_per_step": 1e3 * tp / steps, "path": "one context per replica, in turn"},
            "ratio": vb / vp}


def serve_stream(rate_hz, duration_s, seed=4, budget=200e9, max_cohorts=3, nsample=12, ncheck=4, step_band=1.5):
    """Serving: Poisson arrivals of CG melt jobs (1k-100k beads, 500-5000 steps, each with its own
    corrected table) submitted to the serving loop at their arrival times; every job returns its RDF and mean
    pressure.  Job states are equilibrated (batched, untimed) before the stream starts.  A sample of the jobs
    is then run alone, one context each (the plain path), for the throughput ratio, and the ``ncheck``
    smallest jobs of that sample must equal their served results bit for bit."""
    import queue
    import threading
    import numpy as np
    from ibisco_amd import units, workloads
    from ibisco_amd.engine import Engine, Server

    stream = workloads.serving_stream(rate_hz, duration_s, seed)
    systems = workloads.serving_jobs(stream)
    nj = len(stream)
    print(f"[bench] serving: {nj} jobs over {duration_s:.0f} s, {sum(s.natoms for s in systems)} beads",
          file=sys.stderr, flush=True)
    batch, nb = [], 0
    for k, sy in enumerate(systems + [None]):
        if sy is not None:
            batch.append(sy)
            nb += sy.natoms
        if batch and (sy is None or nb > 2_000_000):
            e = Engine(batch)
            workloads.equilibrate(e, batch, stage_steps=100, final_steps=400)
            for q, b in enumerate(batch):
                s, v = e.state(q)
                b.s, b.v = s[:b.natoms].copy(), units.vel_from_reduced(v)
            e.close()
            print(f"[bench] serving: {k}/{nj} job states equilibrated", file=sys.stderr, flush=True)
            batch, nb = [], 0
    sample = sorted(range(nj), key=lambda k: systems[k].natoms)[::max(1, nj // nsample)][:nsample]
    check = set(sorted(sample, key=lambda k: systems[k].natoms)[:ncheck])
    srv = Server(hbm_budget=budget, max_cohorts=max_cohorts, rdf_nbin=150, rdf_dr=0.01, step_band=step_band)
    results = [None] * nj
    ids = queue.Queue()

    errors = []

    def collect():
        for _ in range(nj):
            k, j = ids.get()
            try:
                results[k] = srv.wait(j)
            except Exception as ex:   # a failed job: counted and reported, the stream goes on
                errors.append(f"job {k}: {ex}")

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for k, (ta, _n, steps, _sd) in enumerate(stream):
        lag = t0 + ta - time.perf_counter()
        if lag > 0:
            time.sleep(lag)
        ids.put((k, srv.submit(systems[k], steps, keep_state=k in check)))
        if (k + 1) % 100 == 0:
            print(f"[bench] serving: {k + 1}/{nj} submitted, {srv.stats()['completed']:.0f} done",
                  file=sys.stderr, flush=True)
    while th.is_alive():
        th.join(timeout=10.0)
        print(f"[bench] serving: {srv.stats()['completed']:.0f}/{nj} done", file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    for msg in errors[:5]:
        print(f"[bench] serving: {msg}", file=sys.stderr, flush=True)
    ok = [r for r in results if r is not None]
    lat = np.array([r["t_done"] - r["t_submit"] for r in ok])
    wait = np.array([r["t_admit"] - r["t_submit"] for r in ok])
    work = sum(r["natoms"] * r["nsteps"] for r in ok)
    span = max(r["t_done"] for r in ok) - min(r["t_submit"] for r in ok)
    # the plain path on a sample of the same jobs, and the bit-for-bit check of the smallest of them
    tp, wp, same = 0.0, 0, True
    for k in sample:
        sy, steps = systems[k], stream[k][2]
        e = Engine(sy)
        e.rdf_enable(150, 0.01)
        e.obs()
        ts = time.perf_counter()
        e.step(steps, graph=True)
        e.obs()
        tp += time.perf_counter() - ts
        wp += sy.natoms * steps
        if k in check and results[k] is not None:
            s, v = e.state()
            r = results[k]
            same = same and bool(np.array_equal(s, r["s"]) and np.array_equal(v, r["v"]) and
                                 np.array_equal(e.rdf(0), r["rdf"]) and
                                 np.array_equal(e.averages(0)[:29], r["sums"][:29]))
        e.close()
    vs, vp = work / span, wp / tp
    print(f"[bench] serving: {vs:.3e} atom-steps/s, p50 {np.percentile(lat, 50):.2f} s, "
          f"p99 {np.percentile(lat, 99):.2f} s; plain sample {vp:.3e}", file=sys.stderr, flush=True)
    pres = [units.pressure_to_bar(r["mean_pressure"]) for r in ok]
    return {"workload": "cg_serving_poisson", "rate_jobs_per_s": rate_hz, "arrival_window_s": duration_s,
            "jobs": nj, "jobs_failed": nj - len(ok), "atom_steps": work,
            "offered_atom_steps_per_s": sum(sy.natoms * st_[2] for sy, st_ in zip(systems, stream)) / duration_s,
            "served": {"value": vs, "span_s": span, "latency_p50_s": float(np.percentile(lat, 50)),
                       "latency_p99_s": float(np.percentile(lat, 99)),
                       "admission_wait_p50_s": float(np.percentile(wait, 50)),
                       "admission_wait_p99_s": float(np.percentile(wait, 99)),
                       "batches": int(st["cohorts"]), "mean_jobs_per_batch": round(st["mean_cohort_jobs"], 2),
                       "peak_concurrent_batches": int(st["peak_running"]),
                       "peak_resident_bytes": int(st["peak_bytes"]), "budget_bytes": int(st["budget"])},
            "plain_sample": {"value": vp, "jobs": len(sample), "path": "one context per job, in turn"},
            "max_concurrent_batches": max_cohorts, "step_band": step_band, "ratio": vs / vp, "alone_equals_served_bitwise": same, "jobs_checked_bitwise": len(check),
            "mean_pressure_bar_p50": float(np.median(pres))}


def hybrid_run(steps, warmup, seed=7):
    """Hybrid melt: 500,000 atoms' worth of chains of 60 (30 % atomistic, one centre-of-mass site per 6 atoms;
    70 % coarse-grained, 10 beads each), three LJ tables with drawn sigma/epsilon, rc 1.4 nm, 1 fs."""
    from ibisco_amd import system, workloads
    from ibisco_amd.engine import Engine
    nch = 500_000 // 60
    nat = int(round(0.3 * nch))
    sy = system.hybrid_melt(nat, nch - nat, chain_len=60, seed=seed, avoid=None, dt=0.001)
    n = sy.natoms + sy.nvs
    print(f"[bench] hybrid: {sy.natoms} atoms+beads, {sy.nvs} sites, box {sy.box[0]:.2f} nm", file=sys.stderr,
          flush=True)
    e = Engine(sy)
    workloads.equilibrate(e, sy, stage_steps=200, final_steps=2000)
    e.step(warmup, graph=True)
    e.obs()
    t0 = time.perf_counter()
    e.step(steps, graph=True)
    ob = e.obs()
    secs = time.perf_counter() - t0
    e.set_timing(True)
    e.step(50)
    tm = e.timing()
    e.close()
    print(f"[bench] hybrid: {steps} steps in {secs:.3f} s", file=sys.stderr, flush=True)
    return {"workload": "hybrid_com_sites_500k", "atoms": int((sy.itype == 0).sum()), "beads": int((sy.itype == 1).sum()),
            "sites": sy.nvs, "items": n, "steps": steps, "value": sy.natoms * steps / secs,
            "unit": "atom-steps/s (atoms and beads; sites are not integrated)", "ms_per_step": 1e3 * secs / steps,
            "step_breakdown_ms": {k: round(tm[k] / tm["nsteps_timed"], 4) for k in
                                  ("nonbonded", "bonded", "integrate", "rebuild", "step")},
            "temperature_K": float(ob["temp"] * 300.0)}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_nonbonded_ua100k_traffic.md")


def measured_traffic(sites):
    """HBM bytes per k_nonbonded launch from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE scaled by a device copy), when it covers this workload."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_json:"):
                    t = json.loads(line.split(":", 1)[1])
                    if t.get("workload") == "ua_melt_100k_plain" and t.get("sites") == sites:
                        return t["bytes"]
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nchains", type=int, default=400)
    ap.add_argument("--nbeads", type=int, default=250)
    ap.add_argument("--eager", action="store_true",
                    help="launch kernel by kernel instead of through HIP graphs (profiling runs)")
    ap.add_argument("--replicas", type=int, default=256, metavar="NREP",
                    help="also compare NREP CG melt replicas batched in one context against one context each "
                         "(0 skips it)")
    ap.add_argument("--replica-steps", type=int, default=200)
    ap.add_argument("--serve-cohorts", type=int, default=3, help="batches the serving loop runs at once")
    ap.add_argument("--serve-band", type=float, default=1.5,
                    help="serving: batch only jobs whose step counts are within this factor (<= 1: off)")
    ap.add_argument("--hybrid-steps", type=int, default=500,
                    help="also time the 500k-atom hybrid system for this many steps (0 skips it)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="also run the serving loop on Poisson arrivals (20 jobs/s) for this many seconds "
                         "(0 skips it)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        n = args.calib_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_read": a.numel() * 4, "calib_copy_bytes_written": b.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); replicas and jobs shard across GPUs "
              "without communication (one serving loop or batch per GPU)", file=sys.stderr)
        sys.exit(2)

    from ibisco_amd import workloads
    from ibisco_amd.engine import Engine

    sysm = workloads.ua_melt_100k(nchains=args.nchains, nbeads=args.nbeads)
    n = sysm.natoms
    print(f"[bench] {n} sites, box {sysm.box[0]:.3f} nm", file=sys.stderr, flush=True)
    e = Engine(sysm)
    workloads.equilibrate(e, sysm, stage_steps=200, final_steps=1000)
    print("[bench] equilibrated", file=sys.stderr, flush=True)
    e.step(args.warmup, graph=not args.eager)
    e.obs()   # synchronise
    t0 = time.perf_counter()
    e.step(args.steps, graph=not args.eager)
    ob = e.obs()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = n * args.steps / secs
    print(f"[bench] {args.steps} steps in {secs:.3f} s, T = {ob['temp'] * 300:.1f} K", file=sys.stderr, flush=True)

    # multiple time step: bonded terms every 1 fs step, nonbonded every 4 (r-RESPA impulse form)
    e.set_respa(4)
    e.step(40, graph=not args.eager)
    e.obs()
    nr = max(400, (args.steps // 10) // 40 * 40)
    t0 = time.perf_counter()
    e.step(nr, graph=not args.eager)
    e.obs()
    tr = time.perf_counter() - t0
    e.set_respa(1)
    respa = {"inner_dt_fs": 1.0, "nonbonded_every": 4, "inner_steps": nr,
             "atom_inner_steps_per_s": n * nr / tr, "atom_outer_steps_per_s": n * nr / 4 / tr,
             "ms_per_outer_step": 1e3 * tr / (nr / 4), "ns_per_day": nr * 1e-6 / tr * 86400,
             "plain_1fs_ns_per_day": args.steps * 1e-6 / secs * 86400}
    # eager timing with the list-skin autotuner off (fixed delta = skin/3 from the next rebuild)
    nt = min(200, max(20, args.steps // 10))
    e.set_skin_tuning(False)
    e.step(10)
    e.set_timing(True)
    e.step(nt)
    tm_fixed = e.timing()
    e.set_timing(False)
    e.set_skin_tuning(True)
    e.step(20)
    # nonbonded kernel: HIP events on the library stream over an eager run (the last kernels of this context,
    # so a kernel trace of this command ends with these launches)
    e.set_timing(True)
    e.step(nt)
    tm = e.timing()
    e.set_timing(False)
    nb_ms = tm["nonbonded"] / tm["nsteps_timed"]
    delta = float(e.skin(0)[0])
    skin_tuning = {"delta_nm": round(delta, 4),
                   "fixed_delta_nm": round((sysm.control.rlist - sysm.control.cutoff) / 3.0, 4),
                   "nb_ms_tuned": round(nb_ms, 5),
                   "nb_ms_fixed": round(tm_fixed["nonbonded"] / tm_fixed["nsteps_timed"], 5)}
    pairs = e.pairs(0, 0)
    prl_n = len(pairs) / n                        # half-list pairs per site
    b_nb = 4.0 * prl_n + 28.0                     # algorithmic bytes per site-step
    achieved = b_nb * n / (nb_ms * 1e-3) / 1e9    # GB/s
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n),
            "kernel": "k_nonbonded", "kernel_ms": round(nb_ms, 5), "timed_launches": int(tm["nsteps_timed"]),
            "bytes_per_site_step": round(b_nb, 1), "half_list_pairs_per_site": round(prl_n, 2)}
    cpu = None
    if not args.no_cpu_baseline:
        from ibisco_amd import units
        pos, vel = e.state()
        sysm.s = pos[:n].copy()
        sysm.v = units.vel_from_reduced(vel)
        print("[bench] timing the reference CPU path", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sysm, args.cpu_steps, args.cpu_threads)
    batched = replicas_compare(args.replicas, args.replica_steps, 50) if args.replicas > 0 else None
    served = serve_stream(20.0, args.serve_seconds, max_cohorts=args.serve_cohorts,
                          step_band=args.serve_band) if args.serve_seconds > 0 else None
    hybrid = hybrid_run(args.hybrid_steps, 50) if args.hybrid_steps > 0 else None
    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random-walk melt, equilibrated on device)",
        "config": {"workload": "ua_melt_100k_plain", "sites": n, "chains": args.nchains, "chain_length": args.nbeads,
                   "rc_nm": 1.0, "rlist_nm": 1.3, "nupdate": 10, "dt_fs": 1.0, "ensemble": "NVE",
                   "time_stepping": "single dt (the reference has no multiple time step for non-hybrid systems)",
                   "path": "one system per launch, " + ("eager launches" if args.eager else
                                                         "HIP graphs of four list intervals"),
                   "precision": "FP32 forces, FP64 integration and reductions"},
        "step_breakdown_ms": {k: round(tm[k] / tm["nsteps_timed"], 5) for k in
                              ("nonbonded", "bonded", "integrate", "rebuild", "step")},
        "list_skin_autotuner": skin_tuning,
        "multiple_time_step": respa,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if batched:
        out["batched_vs_plain"] = batched
    if served:
        out["served_vs_plain"] = served
    if hybrid:
        out["hybrid"] = hybrid
    print(json.dumps(out))


if __name__ == "__main__":
    main()
