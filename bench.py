"""Throughput of the per-step force + integration path on one MI355X.

Headline workload: 256 independent CG melt replicas (log-uniform 2,000-50,000
beads, chain length 20-100, each with its own Gaussian-corrected LJ table,
8 beads/nm^3, rc 1.5 nm, list 1.8 nm rebuilt every 10 steps, 5 fs, NVE, RDF
accumulated on the device at every rebuild), synthetic and equilibrated on the
device before timing, all advanced together by HIP graph replays of one
batched context.  One "step" is one full NEW_LOOP step of every replica.  The
plain path (each replica alone in its own context, one after the other, same
protocol) is timed beside it, and the ratio reported.

The warm-up captures and instantiates every graph the timed region replays:
after the --warmup steps, the --steps steps run once untimed, followed by the
eager steps that return to the same phase of the graph block period, so the
timed --steps steps replay exactly those graphs (the counters prove it).

Prints one JSON line: whole-job atom-steps/s, the nonbonded kernel's
algorithmic-byte roofline fraction (HIP events on the library's stream), the
force/energy/pressure error of the device against the FP64 oracle on sampled
replicas, and the FP32 OpenMP restatement of the reference loop (oracle/
cpu_port.c) timed on this host's cores on a bounded sample of the replicas.
Secondary records: the served Poisson stream, the 100k-site united-atom melt
(with its multiple time step) and the 500k-atom hybrid system.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s on 1 GPU, batched/served vs plain, at ≤1e-5 rel RMS force err"
HBM_PEAK_GBS = 8000.0


def steady_run(e, steps, warmup, graph=True):
    """Warm-up, then the timed ``steps``.  With graphs the warm-up ends with one untimed pass of the same
    ``steps`` and the eager steps that return to its phase in the block period, so the timed pass replays
    graphs captured before it.  Returns (seconds, counters of the timed pass)."""
    e.step(warmup, graph=graph)
    if graph:
        period = e.counters()["period"]
        e.step(steps, graph=True)
        back = (-steps) % period
        if back:
            e.step(back)
    e.obs()   # synchronises
    c0 = e.counters()
    t0 = time.perf_counter()
    e.step(steps, graph=graph)
    e.obs()
    secs = time.perf_counter() - t0
    c1 = e.counters()
    tc = {k: c1[k] - c0[k] for k in ("eager_steps", "graph_steps", "graph_launches", "captures")}
    tc["period"], tc["long_block"] = c1["period"], c1["long_block"]
    return secs, tc


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads for the OpenMP baseline: the CPUs this process may run on, capped by OMP_NUM_THREADS."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(systems, states, target_s=4.0, runs=3):
    """The FP32 OpenMP restatement of the reference loop (oracle/cpu_port.c) on the sampled replicas, run one
    after the other from their equilibrated states; rate = sum of beads x steps / sum of loop seconds, median
    of ``runs`` runs, steps chosen so one run takes about ``target_s`` seconds.  It runs in a child process:
    the OpenMP runtime binds the thread that starts it to a core, and every thread created after that (the
    serving loop's among them) would inherit the binding."""
    import copy
    import subprocess
    import tempfile

    import numpy as np
    from ibisco_amd import system as sysmod, units
    threads = host_threads()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cpu_sample.npz")
        arrays = {"nsys": np.array(len(systems), np.int32)}
        for k, (sy, (s, v)) in enumerate(zip(systems, states)):
            one = copy.copy(sy)
            one.s, one.v = s[:sy.natoms].copy(), units.vel_from_reduced(v)
            for key, val in sysmod.to_arrays(one).items():
                arrays[f"sys{k}_{key}"] = val
        np.savez(path, **arrays)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample", path, "--cpu-seconds",
                            str(target_s), "--cpu-runs", str(runs)], env=env, capture_output=True, text=True,
                           timeout=max(600.0, 20 * target_s * runs))
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads([x for x in r.stdout.splitlines() if x.startswith("{")][-1])


def cpu_sample_run(path, target_s, runs):
    """Child process of cpu_baseline: times the OpenMP restatement on the systems saved in ``path``."""
    import copy

    import numpy as np
    from ibisco_amd import system as sysmod
    from oracle.oracle import CpuPort
    with np.load(path) as z:
        n = int(z["nsys"])
        systems = []
        for k in range(n):
            pre = f"sys{k}_"
            systems.append(sysmod.from_arrays({key[len(pre):]: z[key] for key in z.files if key.startswith(pre)}))

    def one_run(nsteps):
        secs, work = 0.0, 0
        for sy in systems:
            p = CpuPort(copy.copy(sy))
            secs += p.run(nsteps, 0)
            work += sy.natoms * nsteps
            del p
        return secs, work

    secs, work = one_run(10)
    nupdate = systems[0].control.nupdate
    nsteps = max(2 * nupdate, int(round(target_s / max(secs, 1e-6) * 10 / nupdate)) * nupdate)
    nsteps = min(nsteps, 2000)
    rates = []
    for _ in range(runs):
        secs, work = one_run(nsteps)
        rates.append(work / secs)
    rates.sort()
    threads = CpuPort.threads()
    return {"value": rates[len(rates) // 2], "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(systems)} of the replicas ({sum(sy.natoms for sy in systems)} beads) one after the "
                      f"other, {nsteps} steps each from their equilibrated states (list rebuilt every {nupdate}), "
                      f"median of {runs} runs (rates {', '.join(f'{r:.3e}' for r in rates)}); FP32 OpenMP "
                      f"restatement of the reference loop, gcc -O2 -fopenmp, in a child process",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "omp_threads": threads}


def accuracy(e, systems, idx):
    """Device forces, energies and virial of the sampled systems of context ``e`` at their current positions
    against the FP64 oracle on the same positions (each with a fresh list)."""
    import copy
    import numpy as np
    from ibisco_amd import units
    from oracle.oracle import Oracle, nb_virial_magnitude
    e.compute_forces(rebuild=True)
    worst = {"force_rel_rms": 0.0, "energy_rel": 0.0, "virial_rel": 0.0, "virial_rel_of_magnitude": 0.0,
             "pressure_rel": 0.0}
    for k in idx:
        sy = systems[k]
        s, v = e.state(k)
        one = copy.copy(sy)
        one.s, one.v = s[:sy.natoms].copy(), units.vel_from_reduced(v)
        o = Oracle(one)
        o.update_lists()
        o.new_force()
        fo, fg = o.forces(), e.forces(k)
        oo, og = o.obs(), e.obs(k)
        eo = sum(float(oo[q].sum()) for q in ("vnb", "vbond", "vangle", "vtors", "voop"))
        eg = sum(float(og[q].sum()) for q in ("vnb", "vbond", "vangle", "vtors", "voop"))
        po, pg = oo["pt"], og["pt"]
        tr = abs(po[:3].sum())
        worst["force_rel_rms"] = max(worst["force_rel_rms"], float(np.linalg.norm(fg - fo) / np.linalg.norm(fo)))
        worst["energy_rel"] = max(worst["energy_rel"], abs(eg - eo) / abs(eo))
        worst["pressure_rel"] = float(worst["pressure_rel"])
        worst["virial_rel"] = max(worst["virial_rel"], float(np.abs(pg - po).max()) / tr)
        mag = nb_virial_magnitude(o, one)
        if mag:
            worst["virial_rel_of_magnitude"] = max(worst["virial_rel_of_magnitude"],
                                                   float(np.abs(pg - po).max()) / mag)
        worst["pressure_rel"] = max(worst["pressure_rel"], abs(pg[:3].sum() - po[:3].sum()) / tr)
    worst.update(systems_checked=len(idx), reference="FP64 oracle (oracle/ibx_oracle.c) on the same positions",
                 gates={"force_rel_rms": 1e-5, "energy_rel": 1e-7, "virial_rel": 1e-7},
                 virial_scales="virial_rel: max |dPT| / |PT11+PT22+PT33| (the net virial); virial_rel_of_magnitude: "
                               "max |dPT| / sum over in-cutoff pairs of |F| d_k^2 (the terms it sums)")
    return worst


def sample_indices(systems, n):
    """n systems spread over the size range (by bead count)."""
    order = sorted(range(len(systems)), key=lambda k: systems[k].natoms)
    if n >= len(order):
        return order
    if n <= 1:
        return order[len(order) // 2:][:max(n, 0)]   # the median system, or none
    return [order[round(q * (len(order) - 1) / (n - 1))] for q in range(n)]


def tile_record(e):
    """The force launch layout the tile autotuner chose for a context, and what it measured."""
    t, ms = e.tiles()
    names = ["one workgroup per block", "persistent, 5 workgroups per CU, 1 block a take",
             "persistent, 5 per CU, 2 blocks a take", "persistent, 4 per CU, 1 block a take"]
    chosen = names[0] if t[0] == 0 else f"persistent, {t[2]} workgroups per CU, {t[1]} block(s) a take"
    return {"force_launch": chosen, "tiles": t,
            "tuner_ms_per_force_launch": {n: round(m, 4) for n, m in zip(names, ms)}}


def batched_leg(args, seed=3, with_plain=True):
    """Config 3: the replicas in one context (every kernel launch covers all of them), then each alone."""
    import copy
    from ibisco_amd import units, workloads
    from ibisco_amd.engine import Engine

    systems = workloads.cg_replicas(nrep=args.replicas, seed=seed)
    beads = [s.natoms for s in systems]
    total = sum(beads)
    print(f"[bench] {len(systems)} replicas, {total} beads (min {min(beads)}, max {max(beads)})", file=sys.stderr,
          flush=True)
    eb = Engine(systems)
    workloads.equilibrate(eb, systems, stage_steps=100, final_steps=400)
    print("[bench] replicas equilibrated", file=sys.stderr, flush=True)
    eb.rdf_enable(150, 0.01, every=1)
    secs, tc = steady_run(eb, args.steps, args.warmup, graph=not args.eager)
    tiles = tile_record(eb)
    value = total * args.steps / secs
    print(f"[bench] batched: {args.steps} steps in {secs:.4f} s ({value:.3e} atom-steps/s), timed counters {tc}",
          file=sys.stderr, flush=True)
    states = [eb.state(k) for k in range(len(systems))]
    # nonbonded kernel: HIP events on the library stream over eager steps (kernel classes launched apart)
    nt = 20
    eb.set_timing(True)
    eb.step(nt)
    tm = eb.timing()
    eb.set_timing(False)
    nb_ms = tm["nonbonded"] / tm["nsteps_timed"]
    prl = sum(eb.pair_count(k) for k in range(len(systems)))   # half-list pairs of the last rebuild
    idx = sample_indices(systems, args.check)
    acc = accuracy(eb, systems, idx)
    print(f"[bench] batched: accuracy {acc}", file=sys.stderr, flush=True)
    eb.close()
    # the plain path: each replica alone, from the same state, same protocol
    tp = 0.0
    plain_tc = {"captures": 0, "eager_steps": 0, "graph_steps": 0}
    for k, sy in enumerate(systems if with_plain else []):
        one = copy.copy(sy)
        one.s, one.v = states[k][0][:sy.natoms].copy(), units.vel_from_reduced(states[k][1])
        e = Engine(one)
        e.rdf_enable(150, 0.01, every=1)
        t, c = steady_run(e, args.steps, args.warmup, graph=not args.eager)
        tp += t
        for q in plain_tc:
            plain_tc[q] += c[q]
        e.close()
        if (k + 1) % 64 == 0:
            print(f"[bench] plain path: {k + 1}/{len(systems)} replicas", file=sys.stderr, flush=True)
    vp = total * args.steps / tp if tp > 0 else None
    if vp:
        print(f"[bench] replicas: batched {value:.3e}, plain {vp:.3e} atom-steps/s", file=sys.stderr, flush=True)
    return dict(systems=systems, states=states, total=total, secs=secs, value=value, timed_counters=tc, tm=tm,
                nb_ms=nb_ms, prl=prl, accuracy=acc, plain_value=vp, plain_secs=tp, plain_counters=plain_tc,
                check=idx, tiles=tiles)


def serve_stream(rate_hz, duration_s, seed=4, budget=200e9, max_cohorts=3, nsample=12, ncheck=4, step_band=1.5,
                 capacity=True, spare_places=0):
    """Serving: Poisson arrivals of CG melt jobs (1k-100k beads, 500-5000 steps, each with its own
    corrected table) submitted to the serving loop at their arrival times; every job returns its RDF and mean
    pressure.  Job states are equilibrated (batched, untimed) before the stream starts.  A sample of the jobs
    is then run alone, one context each (the plain path), for the throughput ratio, and the ``ncheck``
    smallest jobs of that sample must equal their served results bit for bit.  With ``capacity``, the same
    jobs then arrive three times faster for eight seconds, so that jobs stay queued: the atom-steps run per
    second meanwhile is the loop's capacity."""
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
    srv = Server(hbm_budget=budget, max_cohorts=max_cohorts, rdf_nbin=150, rdf_dr=0.01, step_band=step_band,
                 spare_places=spare_places)
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
    tmg = srv.timing()
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
    cap = serve_capacity(systems, stream, vp, budget, max_cohorts, step_band, rate_hz,
                         spare_places=spare_places) if capacity else None
    return {"workload": "cg_serving_poisson", "rate_jobs_per_s": rate_hz, "arrival_window_s": duration_s,
            "jobs": nj, "jobs_failed": nj - len(ok), "atom_steps": work,
            "offered_atom_steps_per_s": sum(sy.natoms * st_[2] for sy, st_ in zip(systems, stream)) / duration_s,
            "served": {"value": vs, "span_s": span, "latency_p50_s": float(np.percentile(lat, 50)),
                       "latency_p99_s": float(np.percentile(lat, 99)),
                       "admission_wait_p50_s": float(np.percentile(wait, 50)),
                       "admission_wait_p99_s": float(np.percentile(wait, 99)),
                       "batches": int(st["cohorts"]), "mean_jobs_per_batch": round(st["mean_cohort_jobs"], 2),
                       "peak_concurrent_batches": int(st["peak_running"]),
                       "jobs_given_free_places": int(st["admitted_running"]),
                       "peak_resident_bytes": int(st["peak_bytes"]), "budget_bytes": int(st["budget"]),
                       "host_timing": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tmg.items()}},
            "plain_sample": {"value": vp, "jobs": len(sample), "path": "one context per job, in turn"},
            "max_concurrent_batches": max_cohorts, "step_band": step_band,
            "admission": "continuous: queued jobs take free places of running batches at block boundaries",
            "ratio": vs / vp, "offered_vs_plain": (sum(sy.natoms * st_[2] for sy, st_ in zip(systems, stream)) /
                                                   duration_s) / vp,
            "ratio_note": "the served rate is bounded by the offered load: at this arrival rate the ratio cannot "
                          "exceed offered_vs_plain; the capacity record measures the loop under overload",
            "alone_equals_served_bitwise": same, "jobs_checked_bitwise": len(check),
            "mean_pressure_bar_p50": float(np.median(pres)),
            "capacity": cap}


def serve_capacity(systems, stream, vp, budget, max_cohorts, step_band, rate_hz, spare_places=0, factor=3.0,
                   duration_s=8.0,
                   seed=5):
    """The serving loop's capacity: jobs of the stream (cycled) arriving ``factor`` times faster than the stream
    (Poisson) for ``duration_s`` seconds, to a loop set up as for the stream, so that jobs are always queued:
    the atom-steps run per second (every live place of every batch) from t = 2 s to the end of the arrivals.
    The jobs left then run to completion before the loop closes."""
    import numpy as np
    from ibisco_amd.engine import Server

    rng = np.random.default_rng(seed)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / (factor * rate_hz))
        if t >= duration_s:
            break
        arrivals.append(t)
    n = len(systems)
    offered = sum(systems[k % n].natoms * stream[k % n][2] for k in range(len(arrivals))) / duration_s
    srv = Server(hbm_budget=budget, max_cohorts=max_cohorts, rdf_nbin=150, rdf_dr=0.01, step_band=step_band,
                 spare_places=spare_places)
    samples, ids = [], []
    t0 = time.perf_counter()
    nxt = 0.0

    def sample_until(tend):
        nonlocal nxt
        while True:
            now = time.perf_counter() - t0
            if now >= nxt:
                st = srv.stats()
                samples.append((now, st["atom_steps_run"], st["queued"]))
                nxt += 0.25
            if now >= tend:
                return
            time.sleep(max(0.0, min(tend, nxt) - now))

    for k, ta in enumerate(arrivals):
        sample_until(ta)
        ids.append(srv.submit(systems[k % n], stream[k % n][2]))
    sample_until(duration_s)
    print(f"[bench] serving capacity: {len(ids)} jobs submitted over {duration_s:.0f} s, draining",
          file=sys.stderr, flush=True)
    for j in ids:
        srv.wait(j)
    span = time.perf_counter() - t0
    st = srv.stats()
    tmg = srv.timing()
    srv.close()
    win = [x for x in samples if 2.0 <= x[0] <= duration_s]
    rate = (win[-1][1] - win[0][1]) / max(1e-9, win[-1][0] - win[0][0])
    min_queued = int(min(x[2] for x in win))
    total = sum(systems[k % n].natoms * stream[k % n][2] for k in range(len(ids)))
    print(f"[bench] serving capacity: {rate:.3e} atom-steps/s over {win[-1][0] - win[0][0]:.1f} s with at least "
          f"{min_queued} jobs queued; offered {offered:.3e}", file=sys.stderr, flush=True)
    return {"value": rate, "ratio_vs_plain": rate / vp, "window_s": [round(win[0][0], 2), round(win[-1][0], 2)],
            "min_jobs_queued_in_window": min_queued, "rate_jobs_per_s": factor * rate_hz,
            "offered_atom_steps_per_s": offered, "jobs": len(ids), "jobs_failed": int(st["failed"]),
            "batches": int(st["cohorts"]), "jobs_given_free_places": int(st["admitted_running"]),
            "whole_run_atom_steps_per_s": total / span, "whole_run_s": round(span, 2),
            "protocol": f"Poisson arrivals at {factor:g}x the stream's rate for {duration_s:g} s, the stream's jobs "
                        "cycled; atom-steps run per second from t = 2 s to the end of the arrivals",
            "host_timing": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tmg.items()}}


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


def ua_leg(args):
    """Config 1 (secondary record): the 100k-site united-atom melt alone, single 1 fs steps, then bonded 1 fs /
    nonbonded 4 fs (RESPA), with its nonbonded kernel's roofline."""
    from ibisco_amd import workloads
    from ibisco_amd.engine import Engine
    sysm = workloads.ua_melt_100k()
    n = sysm.natoms
    print(f"[bench] UA melt: {n} sites, box {sysm.box[0]:.3f} nm", file=sys.stderr, flush=True)
    e = Engine(sysm)
    workloads.equilibrate(e, sysm, stage_steps=200, final_steps=1000)
    secs, tc = steady_run(e, args.ua_steps, args.warmup, graph=not args.eager)
    tiles = tile_record(e)
    e.set_respa(4)
    nr = max(80, args.ua_steps // 40 * 40)
    tr, _ = steady_run(e, nr, 40, graph=not args.eager)
    e.set_respa(1)
    e.set_timing(True)
    e.step(50)
    tm = e.timing()
    e.set_timing(False)
    nb_ms = tm["nonbonded"] / tm["nsteps_timed"]
    b_nb = 4.0 * e.pair_count() / n + 28.0
    acc = accuracy(e, [sysm], [0])
    e.close()
    ach = b_nb * n / (nb_ms * 1e-3) / 1e9
    print(f"[bench] UA melt: {n * args.ua_steps / secs:.3e} atom-steps/s, NB {nb_ms * 1e3:.1f} us", file=sys.stderr,
          flush=True)
    return {"workload": "ua_melt_100k", "sites": n, "steps": args.ua_steps, "value": n * args.ua_steps / secs,
            "ms_per_step": 1e3 * secs / args.ua_steps, "timed_counters": tc, "tile_autotuner": tiles,
            "step_breakdown_ms": {k: round(tm[k] / tm["nsteps_timed"], 5) for k in
                                  ("nonbonded", "bonded", "integrate", "rebuild", "step")},
            "nonbonded_roofline": {"achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                                   "bytes_per_site_step": round(b_nb, 1), "kernel_ms": round(nb_ms, 5),
                                   "traffic": measured_traffic("ua_melt_100k")},
            "multiple_time_step": {"inner_dt_fs": 1.0, "nonbonded_every": 4, "inner_steps": nr,
                                   "ms_per_outer_step": 1e3 * tr / (nr / 4), "ns_per_day": nr * 1e-6 / tr * 86400,
                                   "plain_1fs_ns_per_day": args.ua_steps * 1e-6 / secs * 86400},
            "accuracy": acc}


TRAFFIC_FILES = {"cg_replicas_256": "k_nonbonded_cg_replicas_traffic.md", "ua_melt_100k": "k_nonbonded_ua100k_traffic.md"}


def measured_traffic(workload):
    """HBM bytes per nonbonded launch from the committed rocprofv3 counter summary of this workload
    (FETCH_SIZE and WRITE_SIZE scaled by a device copy), when there is one."""
    try:
        with open(os.path.join(ROOT, "profiles", TRAFFIC_FILES[workload])) as fh:
            for line in fh:
                if line.startswith("traffic_json:"):
                    t = json.loads(line.split(":", 1)[1])
                    if t.get("workload") == workload:
                        return t["bytes"]
    except (OSError, KeyError, ValueError):
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps of the batched replicas")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--check", type=int, default=8, help="replicas checked against the FP64 oracle")
    ap.add_argument("--legs", default="plain,cpu,served,ua,hybrid",
                    help="secondary legs besides the batched headline: plain, cpu, served, ua, hybrid")
    ap.add_argument("--eager", action="store_true",
                    help="launch kernel by kernel instead of through HIP graphs (profiling runs)")
    ap.add_argument("--ua-steps", type=int, default=2000)
    ap.add_argument("--hybrid-steps", type=int, default=500)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-cohorts", type=int, default=3, help="batches the serving loop runs at once")
    ap.add_argument("--serve-band", type=float, default=1.0,
                    help="serving: batch only jobs whose step counts are within this factor (<= 1: off; off "
                         "measured the higher capacity, with free places refilled as jobs end)")
    ap.add_argument("--serve-spares", type=int, default=0,
                    help="serving: spare places per size class in a new batch (0: the loop's default)")
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="CPU baseline: seconds per run (3 runs)")
    ap.add_argument("--cpu-runs", type=int, default=3, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-sample", default="", help=argparse.SUPPRESS)   # the CPU baseline's child process
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB five times (HBM counter calibration)")
    ap.add_argument("--only-ua", action="store_true",
                    help="only the 100k-site united-atom leg, printed alone (profiling runs)")
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
    if args.cpu_sample:
        print(json.dumps(cpu_sample_run(args.cpu_sample, args.cpu_seconds, args.cpu_runs)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); replicas and jobs shard across GPUs "
              "without communication (one serving loop or batch per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.only_ua:
        print(json.dumps({"ua_melt_100k": ua_leg(args)}))
        return
    legs = set(x for x in args.legs.split(",") if x)
    if "plain" not in legs:
        args_plain = False
    else:
        args_plain = True

    B = batched_leg(args, with_plain=args_plain)
    total, secs = B["total"], B["secs"]
    b_nb = 4.0 * B["prl"] / total + 28.0          # algorithmic bytes per bead-step (half-list int32 entries)
    achieved = b_nb * total / (B["nb_ms"] * 1e-3) / 1e9
    tm = B["tm"]
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("cg_replicas_256"),
            "kernel": "k_nonbonded (the nonbonded role of k_step_forces in graph steps)",
            "kernel_ms": round(B["nb_ms"], 5), "timed_launches": int(tm["nsteps_timed"]),
            "bytes_per_bead_step": round(b_nb, 1), "half_list_pairs_per_bead": round(B["prl"] / total, 2),
            "units_per_launch": total}
    cpu = None
    if "cpu" in legs:
        print("[bench] timing the FP32 OpenMP restatement on the host", file=sys.stderr, flush=True)
        idx = B["check"]
        cpu = cpu_baseline([B["systems"][k] for k in idx], [B["states"][k] for k in idx], target_s=args.cpu_seconds)
        cpu["accuracy_of_gpu_path"] = B["accuracy"]
    out = {
        "metric": METRIC, "value": B["value"], "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random-walk CG melts with seeded Gaussian-corrected tables, equilibrated on device)",
        "config": {"workload": "cg_replicas_256_batched", "replicas": len(B["systems"]), "beads": total,
                   "beads_per_replica": "log-uniform 2,000-50,000, chain length 20-100", "density_per_nm3": 8.0,
                   "rc_nm": 1.5, "rlist_nm": 1.8, "nupdate": 10, "dt_fs": 5.0, "ensemble": "NVE",
                   "rdf": "150 bins of 0.01 nm, every rebuild",
                   "path": "one batched context, " + ("eager launches" if args.eager else
                                                      "HIP graph replays captured in the warm-up"),
                   "precision": "FP32 forces, FP64 integration and reductions"},
        "timed_counters": B["timed_counters"],
        "tile_autotuner": B["tiles"],
        "step_breakdown_ms": {k: round(tm[k] / tm["nsteps_timed"], 5) for k in
                              ("nonbonded", "bonded", "integrate", "rebuild", "step")},
        "accuracy": B["accuracy"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if args_plain:
        out["batched_vs_plain"] = {"batched": B["value"], "plain": B["plain_value"],
                                   "ratio": B["value"] / B["plain_value"],
                                   "plain_ms_per_step": 1e3 * B["plain_secs"] / args.steps,
                                   "plain_path": "each replica alone in its own context, one after the other, "
                                                 "same warm-up and graph protocol",
                                   "plain_timed_counters": B["plain_counters"]}
    if "served" in legs and args.serve_seconds > 0:
        out["served_vs_plain"] = serve_stream(20.0, args.serve_seconds, max_cohorts=args.serve_cohorts,
                                              step_band=args.serve_band, spare_places=args.serve_spares)
    if "ua" in legs:
        out["ua_melt_100k"] = ua_leg(args)
    if "hybrid" in legs and args.hybrid_steps > 0:
        out["hybrid"] = hybrid_run(args.hybrid_steps, 50)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
