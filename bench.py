"""Benchmark: optimal-GST queries/sec on one MI355X, batched vs plain.

Workload: the 1.1M-vertex R-MAT graph (a=0.57, b=c=0.19, ~3.0M undirected
edges, int32 weights uniform in [1,100]) with 500,000 groups (geometric sizes,
mean 6), queries of k = 5 distinct group ids, no diameter bound.  The plain
path answers queries one call at a time (gst_solve); the batched path hands
`--batch` queries per step to gst_solve_batch, which runs them together in
shared engine steps: one group-distance job per up to 256 labels (all the
batch's groups in one multi-label SSSP) and the per-query DP on the live set.

Prints one JSON line.  ``value`` is the batched throughput; the plain
throughput, latency percentiles, parity checks (plain vs batched costs, tree
validation, the CPU oracle on a bounded sample), the roofline of the dominant
step kernel and the CPU baseline ride along in the same object.

Schedule: the batched engine's options (work-item size, grid, hipGraph
replay, admitted slots) come from the autotuner's results file when it holds
this graph and query shape (``--tune`` re-tunes on a sample batch first).

Roofline: one extra batch (not in the timed steps) runs with every step's
lean and heavy kernels bracketed by HIP events on the engine stream; each
kernel's average launch time and the algorithmic bytes its items moved (from
the device counters, models below) give its `achieved`; the dominant one is
`roofline`, both are under `roofline.kernels`, with the measured HBM bytes and
L2 atomics per launch from the committed rocprofv3 summary.

After the headline legs (plain, batched, roofline, serving on the same graph)
come the other configurations on one GPU: k = 5 with hop bound D = 4
(`configs.d4`), 4,096 mixed-k queries in one batch (`configs.kmix_4096`), the
CPU baseline on the host cores, and the 3M-vertex serving configuration at
three offered rates (`serving_c5`).  `--no-extra` skips the extra legs.
"""
import argparse
import json
import sys
import time

import numpy as np

METRIC = "optimal-GST queries/sec + p50/p99 latency, batched vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else None


def job_fill_bytes(g, V, ks, ND=1):
    """Algorithmic bytes of the group-distance jobs' FILL: 4 B per (vertex,
    label) of height 0 + 24 B per vertex (3 masks).  An unbounded job with hub
    bounds also reads each vertex's row of hub distances once (2 B per hub in
    the packed 16-bit fill, else 4 B)."""
    labels = int(sum(ks))
    jobs = max(1, g.get_option("batch_jobs"))
    lp = labels + (chunk_labels(ND) - 1) * jobs // 2  # rows padded to whole chunks (half a chunk on average)
    b = 4 * V * lp + 24 * V * jobs
    if hub_fill(g, ND):
        b += g.get_option("hub_bytes") * HUBS * V * jobs
    return b


HUBS = 64  # hub vertices of the shortest-path backbone build (kBbHubs)
JOB_CHUNK_LABELS = 16  # labels per dirty chunk of an unbounded job's row (kJobChunkFlat)
JOB_CHUNK_LABELS_HOP = 4  # labels per dirty chunk of a hop job's rows (kJobChunkHop)


def chunk_labels(ND):
    return JOB_CHUNK_LABELS if ND == 1 else JOB_CHUNK_LABELS_HOP


def hub_fill(g, ND):
    """Unbounded jobs start their rows at hub bounds: their FILL runs in the
    heavy kernel (it keeps 64 hub bounds per lane in registers)."""
    return ND == 1 and g.get_option("hub_bounds") != 0 and g.get_option("backbone") != 0


def lean_bytes(g, res_list, V, ks, ND=1):
    """Algorithmic bytes of the lean kernel's items for these queries (one
    group-distance job per up to 256 labels; ND = D + 1 rows per vertex for a
    hop job, 1 otherwise):
      job FILL        (job_fill_bytes; in the heavy kernel for unbounded jobs
                      with hub bounds)
      job ROUND       per frontier entry 24 B (list id, arc offset, mask, row
                      pointer) + 12 B clear, and its dirty chunks of the source
                      row (CL labels of 4 B each, read once per entry); per arc
                      8 B (column + weight); per label relaxed 4 B (target
                      chunk read); per improved chunk 4 * CL B (write back);
                      CL = 16 labels per chunk for an unbounded job, 4 for a
                      hop job
      job COPY        a hop job's heights 1..D: 4 B read + 4 B written per
                      (vertex, label, height)
      job STAR        4 B per (vertex, label) read, from the row of height
                      floor(D/2) (0 unbounded)
      query FILL      1/4 B per vertex (live and hold bits); a query fed by a
                      job clears only its live vertices' masks, in LIVE"""
    labels = int(sum(ks))
    jobs = max(1, g.get_option("batch_jobs"))
    lp = labels + (chunk_labels(ND) - 1) * jobs // 2
    ent = g.get_option("batch_job_entries")
    arcs = g.get_option("batch_job_arcs")
    relax = g.get_option("batch_job_relaxations")
    impr = g.get_option("batch_job_improvements")
    src = g.get_option("batch_job_src_chunks")
    cl = chunk_labels(ND)
    b = 0 if hub_fill(g, ND) else job_fill_bytes(g, V, ks, ND)
    b += 36 * ent + 8 * arcs + 4 * relax + 4 * cl * (impr + src)
    b += 8 * V * lp * (ND - 1)
    b += 4 * V * lp
    b += (V // 4) * len(res_list)
    return b


def heavy_bytes(g, res_list, V, ks, ND=1):
    """Algorithmic bytes of the heavy kernel's items for a batch of job-fed
    queries: the jobs' live test (each job's rows of height D read once, 4 B
    per vertex and label) and, per query, its LIVE (its live bits, V / 8 B,
    then per live vertex 4 B of list, 24 B of dirty masks cleared and its T
    segment written once, 8 B per state; NS = ceil(D/2) + 1 = ND // 2 + 1
    state heights for a hop query, 1 otherwise), the level-2 merge reading
    the singletons (8 B per vertex, group and state height) and C rows (4 B
    per vertex, group and row) of every live vertex, the later merges and
    FINAL reading the T segments of the active vertices (8 B per state and
    pass), and the level >= 2 grow rounds (24 B per frontier entry, 8 B per
    arc, 16 B per relaxation candidate, 8 B per improvement) -- the counters
    of a query fed by a job hold only these rounds; plus the FILL of
    unbounded jobs with hub bounds (job_fill_bytes)."""
    jobs = max(1, g.get_option("batch_jobs"))
    total = 4 * V * (int(sum(ks)) + (chunk_labels(ND) - 1) * jobs // 2)
    if hub_fill(g, ND):
        total += job_fill_bytes(g, V, ks, ND)
    for r, k in zip(res_list, ks):
        W = 1 << k
        c = r.counters
        nl = c.get("live_vertices", 0)
        na = c.get("active_vertices", 0)
        passes = max(k - 3, 0) + 1  # merges of levels 3..k-1, then FINAL
        NS = ND // 2 + 1
        total += V // 8 + nl * (W * NS * 8 + 28) + nl * k * (8 * NS + 4 * ND) + na * W * NS * 8 * passes
        total += 24 * c["frontier_entries"] + 8 * c["arcs"] + 16 * c["relaxations"] + 8 * c["improvements"]
    return total


PROFILE = "step_kernels_rmat1.1M_k5_batch256.md"
# MI355X L2 atomic rate used as the atomics roofline: 8 XCDs x 16 L2 channels
# x 1 atomic per channel per clock x 2.4 GHz (an estimate from the cache
# organisation, not a published figure)
L2_ATOMIC_PEAK = 8 * 16 * 2.4e9


def profiled(kernel, what):
    """Per-launch figure of `kernel` on the headline workload from the committed
    rocprofv3 summary: what = "traffic_bytes_per_launch" (FETCH_SIZE x2.0 +
    WRITE_SIZE x1.0, the copy calibration) or "l2_atomics_per_launch"
    (TCC_ATOMIC_sum), or None."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", PROFILE)
    try:
        for line in open(path):
            if line.startswith(f"{what} {kernel}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def poisson_serve(g, qs, Ds, rate, seed):
    """Open-loop Poisson arrivals of the requests (qs, Ds) into a fresh server
    at `rate` per second; returns (achieved q/s, p50 ms, p99 ms, results)."""
    import threading
    from gpugst_amd import Server
    n = len(qs)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, n)
    srv = Server(g)
    tickets = [None] * n

    def submitter():
        t_next = time.perf_counter()
        for i in range(n):
            t_next += gaps[i]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[i] = srv.submit(qs[i], D=Ds[i])

    ts = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    res = [srv.wait(t) for t in tickets]
    span = time.perf_counter() - ts
    steps, busy_us = srv.stat("steps"), srv.stat("busy_us")
    srv.close()
    lat = [r.latency_ms for r in res]
    dev = [r.device_ms for r in res]
    wait = [r.latency_ms - r.device_ms for r in res]
    detail = {"device_p50_ms": round(pct(dev, 50), 2), "queue_wait_p50_ms": round(pct(wait, 50), 2),
              "engine_steps": steps, "device_busy_frac": round(busy_us * 1e-3 / (span * 1e3), 3),
              "avg_step_us": round(busy_us / max(1, steps), 1)}
    return n / span, pct(lat, 50), pct(lat, 99), res, detail


def hit_rate(res, ks):
    """Share of the requests' groups served from the group-distance cache."""
    tot = sum(ks)
    return round(sum(r.counters.get("cached_labels", 0) for r in res) / tot, 4) if tot else None


def serve_config5(log, n_req=96, cache_mb=4096):
    """The serving configuration: Chung-Lu graph, V = 3.0M, 117M undirected
    edges (exponent 2.0, weights in [1, 1000]), 2M groups (geometric sizes,
    mean 10, max 5,000); requests with k uniform in {3..6}, 20% with hop bound
    D = 3.  Capacity = one batch of n_req requests; then n_req Poisson
    requests at each of 0.5x, 0.8x and 1.1x of it, the first 16 answers of
    the first rate checked against one-query calls.  The group-distance cache
    is on (cache_mb); its hit rate is reported per rate."""
    from gpugst_amd import Graph, data
    t0 = time.time()
    inst = data.make_config(5, nq=n_req * 5)
    log(f"serving config: V={inst.V} arcs={inst.col.shape[0]} groups={inst.gptr.shape[0]-1} "
        f"gen {time.time()-t0:.1f}s")
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    g.set_option("cache_mb", cache_mb)
    qs, Ds = list(inst.queries), list(inst.D)
    g.solve_batch(qs[:16], Ds[:16])  # warm-up
    tb = time.perf_counter()
    rb = g.solve_batch(qs[16:16 + n_req], Ds[16:16 + n_req])
    cap = n_req / (time.perf_counter() - tb)
    log(f"serving config: batched capacity {cap:.1f} q/s")
    g.set_option("budget_mb", 0)  # the batch engine hands its arena back to the server
    out = {"workload": "chunglu-3M-117M-edges, k in {3..6}, 20% D=3", "batched_capacity_qps": round(cap, 1),
           "requests_per_rate": n_req, "cache_mb": cache_mb, "rates": []}
    mism = 0
    for i, f in enumerate((0.5, 0.8, 1.1)):
        a = 16 + n_req * (1 + i)
        q, d = qs[a:a + n_req], Ds[a:a + n_req]
        ach, p50, p99, res, det = poisson_serve(g, q, d, f * cap, 21 + i)
        if i == 0:
            for j in range(16):
                mism += int(g.solve(q[j], D=d[j]).cost != res[j].cost)
        out["rates"].append({"offered_qps": round(f * cap, 1), "achieved_qps": round(ach, 1),
                             "p50_ms": round(p50, 1), "p99_ms": round(p99, 1),
                             "cache_hit_rate": hit_rate(res, [len(x) for x in q]), **det})
        log(f"serving config: offered {f*cap:.1f} q/s -> {ach:.1f} q/s, p50 {p50:.1f} ms p99 {p99:.1f} ms")
    out["mix_vs_alone_mismatches"] = mism
    out["batch_infeasible"] = sum(1 for r in rb if not r.feasible)
    g.close()
    return out


def hop_leg(g, inst, log, batch=128, steps=2, plain=16):
    """The diameter-bounded configuration on the same graph: k = 5, D = 4,
    batches of `batch` queries (one untimed warm-up batch), plain one call at
    a time; batched costs equal the plain ones on the shared queries."""
    from gpugst_amd import data
    n = batch * (steps + 1)
    qs = data.make_queries(n, 5, inst.meta["eligible"], 3000 + 4 + 10 * 5 + 3)
    Ds = [4] * n
    tp = time.perf_counter()
    pr = [g.solve(q, D=4) for q in qs[:plain]]
    plain_qps = plain / (time.perf_counter() - tp)
    g.solve_batch(list(qs[:batch]), Ds[:batch])
    tb = time.perf_counter()
    res = []
    for s in range(steps):
        a = batch * (1 + s)
        res += g.solve_batch(list(qs[a:a + batch]), Ds[a:a + batch])
    bq = batch * steps / (time.perf_counter() - tb)
    chk = g.solve_batch(list(qs[:plain]), Ds[:plain])
    mism = sum(int(a.cost != b.cost) for a, b in zip(pr, chk))
    log(f"D=4 leg: batched {bq:.1f} q/s, plain {plain_qps:.1f} q/s, mismatches {mism}")
    return {"workload": "rmat-1.1M k=5 D=4", "batch": batch, "batched_qps": round(bq, 1),
            "plain_qps": round(plain_qps, 1), "batched_vs_plain": round(bq / plain_qps, 2),
            "plain_vs_batched_mismatches": mism, "infeasible": sum(1 for r in res if not r.feasible)}


def kmix_leg(g, inst, log, n=4096, check=32):
    """The batched persistent configuration: 4,096 queries with k uniform in
    {3..6}, unbounded, in one gst_solve_batch (continuous admission inside the
    engine); the first `check` answers against one call each."""
    from gpugst_amd import data
    ks = data.uniform_ints(n, 3, 6, 4004)
    qs = [data.make_queries(1, int(ks[i]), inst.meta["eligible"], 4004 * 1_000_003 + i)[0] for i in range(n)]
    t = time.perf_counter()
    res = g.solve_batch(qs)
    dt = time.perf_counter() - t
    dev = [r.device_ms for r in res]
    mism = sum(int(g.solve(qs[i]).cost != res[i].cost) for i in range(check))
    log(f"k-mix leg: {n} queries in {dt:.2f}s -> {n/dt:.1f} q/s, mismatches {mism}")
    return {"workload": "rmat-1.1M k in {3..6} unbounded, one batch of 4096", "queries": n,
            "qps": round(n / dt, 1), "device_p50_ms": round(pct(dev, 50), 1),
            "device_p99_ms": round(pct(dev, 99), 1), "plain_vs_batched_mismatches": mism, "checked": check}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(inst, queries, gpu_res, D, seconds, log):
    """The CPU baseline on the host cores: the oracle's exact solver (pruned
    best-first DPBF for unbounded queries, the hop DP otherwise) over a bounded
    sample of the same queries -- one thread for `seconds` (per-query
    latency, and the reference's convention of timing the main DP loop only,
    CPUNONHOP.h:634-665), then one query per thread on up to 16 threads (the
    box's CPU share) for about as long.  Every answer is compared with the
    GPU's."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    oracle.lib()

    def one(i):
        groups = [inst.group(int(x)) for x in queries[i]]
        t = time.perf_counter()
        if D < 0:
            c, st = oracle.dpbf_pruned(inst.rowptr, inst.col, inst.wt, groups)
            loop = st[3] / 1e9
        else:
            c = oracle.dpbf(inst.rowptr, inst.col, inst.wt, groups, D=D)[0]
            loop = None
        return c, time.perf_counter() - t, loop

    lat, loops, mism, n = [], [], 0, 0
    tc = time.perf_counter()
    while n < len(gpu_res) and time.perf_counter() - tc < seconds:
        c, dt, lp = one(n)
        mism += int(c != gpu_res[n].cost)
        lat.append(dt)
        if lp is not None:
            loops.append(lp)
        n += 1
    single = n / (time.perf_counter() - tc)
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    m = min(len(gpu_res), max(threads, int(single * seconds * threads)))
    tm = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        outs = list(ex.map(one, range(m)))
    multi = m / (time.perf_counter() - tm)
    mism += sum(int(o[0] != gpu_res[i].cost) for i, o in enumerate(outs))
    algo = ("exact pruned best-first DPBF (orc_dpbf_pruned, the exact-mode PrunedDP++ restatement)"
            if D < 0 else "exact hop-bounded DP (orc_dpbf, the HOP_cpu restatement)")
    log(f"cpu baseline: 1 thread {single:.2f} q/s ({n} queries), {threads} threads {multi:.2f} q/s "
        f"({m} queries), mismatches {mism}")
    return {"value": round(multi, 3), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{algo} in oracle/gst_oracle.c on the same queries: first {n} on one thread "
                      f"({sum(lat):.1f} s), then the first {m} one per thread on {threads} threads",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "single_thread_qps": round(single, 3),
            "single_thread_p50_ms": round(pct([x * 1e3 for x in lat], 50), 1),
            "single_thread_p99_ms": round(pct([x * 1e3 for x in lat], 99), 1),
            "main_loop_only_qps": round(len(loops) / sum(loops), 3) if loops and sum(loops) > 0 else None,
            "oracle_checked": n + m, "oracle_mismatches": mism}


def roofline_pass(g, qs, Ds, V, D, log):
    """One further batch with HIP events around every step kernel on the
    engine stream: average launch time of the lean and heavy kernels, their
    algorithmic bytes per launch (device counters, byte models above), and
    the job rounds' atomics per launch; the rocprofv3 summary in profiles/
    gives the measured HBM bytes and L2 atomics of the same workload."""
    g.set_option("reset_stats", 0)
    use_graph = g.get_option("graph")
    g.set_option("graph", 0)  # events bracket eager launches
    g.set_option("time_steps", 1)
    rr = g.solve_batch(list(qs), Ds)
    g.set_option("time_steps", 0)
    g.set_option("graph", use_graph)
    n_steps = max(1, g.get_option("batch_timed_steps"))
    ND = 1 if D < 0 else D + 1
    rk = [len(q) for q in qs]
    kern = {}
    for name, ms, byts in (("gst_lean_kernel", g.get_option("batch_lean_ns") / 1e6 / n_steps, lean_bytes(g, rr, V, rk, ND)),
                           ("gst_heavy_kernel", g.get_option("batch_heavy_ns") / 1e6 / n_steps, heavy_bytes(g, rr, V, rk, ND))):
        ach = byts / n_steps / (ms * 1e-3) / 1e9 if (byts and ms > 0) else 0.0
        kern[name] = {"avg_launch_ms": round(ms, 4), "algorithmic_bytes_per_launch": int(byts / n_steps),
                      "achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}
    job_atomics = g.get_option("batch_job_atomics") / n_steps
    lean = kern["gst_lean_kernel"]
    lean["job_atomics_per_launch"] = int(job_atomics)
    lean["job_atomics_per_s"] = round(job_atomics / (lean["avg_launch_ms"] * 1e-3), 1) if lean["avg_launch_ms"] else None
    log(f"roofline pass: {n_steps} steps; " + "; ".join(
        f"{k} {v['avg_launch_ms']:.3f} ms/launch {v['achieved']:.0f} GB/s" for k, v in kern.items()))
    log("job counters: " + " ".join(f"{f}={g.get_option('batch_job_' + f)}" for f in
                                    ("entries", "arcs", "relaxations", "improvements", "atomics")))
    return kern, int(n_steps), len(rr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256, help="queries per step")
    ap.add_argument("--plain", type=int, default=32, help="queries answered one at a time")
    ap.add_argument("--k", type=int, default=5)
    ap.add_argument("--kmix", action="store_true",
                    help="k drawn uniformly from {3,4,5,6} per query (the mixed-k batched workload)")
    ap.add_argument("--D", type=int, default=-1)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the graph (tests only)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline budget per leg (s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the D=4, 4096 mixed-k and 3M-vertex serving legs")
    ap.add_argument("--max-slots", type=int, default=0)
    ap.add_argument("--serve", type=int, default=512, help="queries of the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="offered queries/s of the serving leg (default: half the batched throughput)")
    ap.add_argument("--cache-mb", type=int, default=4096, help="group-distance cache of the serving legs")
    ap.add_argument("--tune", action="store_true",
                    help="run the schedule autotuner on a sample batch before the batched leg")
    ap.add_argument("--tune-db", default=None,
                    help="autotuner results file (default: gpugst_amd/tuned.json)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "queries are independent and would shard by replicating the graph per GPU.",
              file=sys.stderr)
        sys.exit(2)

    from gpugst_amd import Graph, data, validate_tree

    t0 = time.time()
    inst = data.make_config(3, scale=args.scale, nq=1)
    V = inst.V
    n_roof = 0 if args.no_roofline else args.batch
    n_batched = (args.warmup + args.steps) * args.batch + n_roof
    nq = max(n_batched, args.plain)
    if args.kmix:
        # the mixed-k workload's query stream: k uniform in {3..6}, one seeded draw per query
        ks = data.uniform_ints(nq, 3, 6, 4004)
        queries = [data.make_queries(1, int(ks[i]), inst.meta["eligible"], 4004 * 1_000_003 + i)[0]
                   for i in range(nq)]
    else:
        queries = data.make_queries(nq, args.k, inst.meta["eligible"], 3000 + 4 + 10 * args.k + 1)
    kdesc = "k in {3..6}" if args.kmix else f"k={args.k}"
    log(f"graph V={V} arcs={inst.col.shape[0]} groups={inst.gptr.shape[0]-1} "
        f"queries={nq} {kdesc} D={args.D} gen {time.time()-t0:.1f}s")

    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    if args.max_slots:
        g.set_option("max_slots", args.max_slots)
    Ds = [args.D] * nq

    # ---------------- plain: one query per call
    plain_res, plain_lat = [], []
    plain_qps = None
    if args.plain > 0:
        g.solve(queries[0], D=args.D)  # warm-up (allocates the single-query arena)
        g.set_option("reset_stats", 0)
        tp = time.perf_counter()
        for i in range(args.plain):
            t = time.perf_counter()
            plain_res.append(g.solve(queries[i], D=args.D))
            plain_lat.append((time.perf_counter() - t) * 1e3)
        plain_s = time.perf_counter() - tp
        plain_qps = args.plain / plain_s
        log(f"plain: {args.plain} queries {plain_s:.3f}s -> {plain_qps:.1f} q/s "
            f"p50 {pct(plain_lat, 50):.2f} ms p99 {pct(plain_lat, 99):.2f} ms")

    # ---------------- schedule of the batched engine: the autotuner's choice
    from gpugst_amd import autotune
    db = args.tune_db or autotune.DEFAULT_DB
    kb = "mix" if args.kmix else args.k
    if args.tune:
        sample = list(queries[:args.batch])
        schedule = autotune.tune(g, sample, D=args.D, path=db, log=log, kkey=kb)["best"]
    else:
        schedule = autotune.apply(g, kb, args.D, path=db)
    if args.max_slots:
        g.set_option("max_slots", args.max_slots)
    log(f"schedule: {schedule if schedule else 'engine defaults (no tuned entry for this graph and shape)'}")

    # ---------------- batched: warmup + timed steps
    bres = []
    for w in range(args.warmup):
        qs = queries[w * args.batch:(w + 1) * args.batch]
        bres += g.solve_batch(list(qs), Ds[:len(qs)])
        log(f"warmup step {w} done")
    g.set_option("reset_stats", 0)
    timed = []
    step_ms = []
    tb = time.perf_counter()
    for s in range(args.steps):
        a = (args.warmup + s) * args.batch
        qs = queries[a:a + args.batch]
        t = time.perf_counter()
        r = g.solve_batch(list(qs), Ds[:len(qs)])
        step_ms.append((time.perf_counter() - t) * 1e3)
        timed += r
        log(f"step {s}: {step_ms[-1]:.1f} ms")
    batch_s = time.perf_counter() - tb
    batch_qps = len(timed) / batch_s
    kern_ns = g.get_option("batch_kernel_ns")
    launches = g.get_option("batch_launches")
    log(f"batched: {len(timed)} queries {batch_s:.3f}s -> {batch_qps:.1f} q/s; "
        f"device {kern_ns/1e6:.1f} ms in {launches} epochs")

    # ---------------- roofline pass: one more batch with per-kernel event timing
    roof = None
    if n_roof:
        a = (args.warmup + args.steps) * args.batch
        kern, n_steps, n_q = roofline_pass(g, queries[a:a + args.batch], Ds[:args.batch], V, args.D, log)
        headline = args.D < 0 and args.batch == 256 and args.k == 5 and not args.kmix and args.scale == 1.0
        for name, kv in kern.items():
            kv["traffic"] = profiled(name, "traffic_bytes_per_launch") if headline else None
            at = profiled(name, "l2_atomics_per_launch") if headline else None
            kv["l2_atomics_per_launch"] = at
            kv["l2_atomics_per_s"] = round(at / (kv["avg_launch_ms"] * 1e-3), 1) if at else None
        dom = max(kern, key=lambda k: kern[k]["avg_launch_ms"])
        d = kern[dom]
        roof = {"bound": "hbm", "achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["frac"], "traffic": d["traffic"], "kernel": dom,
                "algorithmic_bytes_per_launch": d["algorithmic_bytes_per_launch"],
                "avg_launch_ms": d["avg_launch_ms"], "launches": n_steps,
                "atomics": {"kernel": dom, "l2_atomics_per_s": d.get("l2_atomics_per_s"),
                            "peak_per_s": L2_ATOMIC_PEAK,
                            "frac": round(d["l2_atomics_per_s"] / L2_ATOMIC_PEAK, 4)
                            if d.get("l2_atomics_per_s") else None,
                            "source": f"profiles/{PROFILE} (rocprofv3 TCC_ATOMIC_sum per launch, same "
                                      "workload) over the live launch time; peak = 8 XCDs x 16 L2 "
                                      "channels x 2.4 GHz, an estimate"},
                "kernels": kern,
                "traffic_source": f"profiles/{PROFILE} (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per "
                                  "launch, same workload)" if d["traffic"] else None,
                "pass": f"{n_q} further queries of the same workload, events around every step kernel"}

    # ---------------- serving: open-loop Poisson arrivals into the server
    serving = None
    if args.serve > 0 and args.D < 0:
        rate = args.serve_rate if args.serve_rate > 0 else 0.5 * batch_qps
        sq = data.make_queries(args.serve, args.k, inst.meta["eligible"], 3000 + 4 + 10 * args.k + 7)
        g.set_option("cache_mb", args.cache_mb)
        ach, p50, p99, sres, det = poisson_serve(g, list(sq), [-1] * args.serve, rate, 11)
        g.set_option("cache_mb", 0)
        serving = {"offered_qps": round(rate, 1), "queries": args.serve, "achieved_qps": round(ach, 1),
                   "p50_ms": round(p50, 2), "p99_ms": round(p99, 2),
                   "p50_vs_plain_p50": round(p50 / pct(plain_lat, 50), 2) if plain_lat else None,
                   "cache_mb": args.cache_mb, "cache_hit_rate": hit_rate(sres, [len(q) for q in sq]),
                   **det, "admission": {k: g.get_option(k) for k in ("serve_chunk", "serve_depth",
                                                                      "serve_wait_us", "serve_batch")},
                   "arrivals": "Poisson (seeded), open loop, gst_server_* continuous stepping: requests "
                               "join the running engine between chunks of steps"}
        log(f"serving: offered {rate:.0f} q/s -> {ach:.0f} q/s, p50 {p50:.2f} ms p99 {p99:.2f} ms, "
            f"cache hit rate {serving['cache_hit_rate']}, {det}")
        # the same requests at half that rate
        ach2, p502, p992, _, det2 = poisson_serve(g, list(sq), [-1] * args.serve, 0.5 * rate, 12)
        serving["at_half_rate"] = {"offered_qps": round(0.5 * rate, 1), "achieved_qps": round(ach2, 1),
                                   "p50_ms": round(p502, 2), "p99_ms": round(p992, 2),
                                   "p50_vs_plain_p50": round(p502 / pct(plain_lat, 50), 2) if plain_lat else None,
                                   **det2}
        log(f"serving: offered {0.5 * rate:.0f} q/s -> {ach2:.0f} q/s, p50 {p502:.2f} ms p99 {p992:.2f} ms, {det2}")

    # ---------------- parity: plain vs batched on shared queries, trees valid
    all_b = bres + timed
    mism = 0
    for i in range(min(len(plain_res), len(all_b))):
        if plain_res[i].cost != all_b[i].cost:
            mism += 1
    bad_trees = 0
    for i, r in enumerate(all_b[:64]):
        if r.feasible:
            v = validate_tree(inst.rowptr, inst.col, inst.wt,
                              [inst.group(int(x)) for x in queries[i]], r.cost, r.root, r.edges, D=args.D)
            bad_trees += 0 if v["ok"] else 1
    infeas = sum(1 for r in timed if not r.feasible)

    # ---------------- the other configurations of the workload, on one GPU
    configs = {}
    if not args.no_extra and args.scale == 1.0 and not args.kmix and args.D < 0:
        configs["d4"] = hop_leg(g, inst, log)
        configs["kmix_4096"] = kmix_leg(g, inst, log)

    # ---------------- CPU baseline on the host cores (also parity)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_leg(inst, queries, all_b, args.D, args.cpu_seconds, log)

    # ---------------- serving on the 3M-vertex serving configuration (the
    # first graph's engines give their device memory back first)
    serving_c5 = None
    if not args.no_extra and args.scale == 1.0:
        g.close()
        serving_c5 = serve_config5(log, cache_mb=args.cache_mb)

    lat = [r.latency_ms for r in timed]
    dev = [r.device_ms for r in timed]
    out = {
        "metric": METRIC,
        "value": round(batch_qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_s * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"rmat-1.1M {kdesc} {'unbounded' if args.D < 0 else f'D={args.D}'} "
                               f"batched, {args.batch} queries/step",
                   "V": V, "arcs": int(inst.col.shape[0]), "groups": int(inst.gptr.shape[0] - 1),
                   "k": "3..6" if args.kmix else args.k, "D": args.D, "batch": args.batch, "scale": args.scale,
                   "schedule": schedule},
        "plain_qps": round(plain_qps, 2) if plain_qps else None,
        "batched_vs_plain": round(batch_qps / plain_qps, 2) if plain_qps else None,
        "latency_ms": {"plain_p50": pct(plain_lat, 50), "plain_p99": pct(plain_lat, 99),
                       "batched_device_p50": pct(dev, 50), "batched_device_p99": pct(dev, 99),
                       "batched_submit_p50": pct(lat, 50), "batched_submit_p99": pct(lat, 99)},
        "parity": {"plain_vs_batched_mismatches": mism, "invalid_trees": bad_trees,
                   "oracle_checked": cpu["oracle_checked"] if cpu else 0,
                   "oracle_mismatches": cpu["oracle_mismatches"] if cpu else 0,
                   "infeasible": infeas},
        "serving": serving,
        "configs": configs,
        "serving_c5": serving_c5,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if serving_c5 is None:
        g.close()


if __name__ == "__main__":
    main()
