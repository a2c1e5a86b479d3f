"""Engine checks run against the CPU build of libgst (test infrastructure).

Started by tests/test_engine_cpu.py in its own process with GST_LIB pointing
at tests/cpu_sim/_build/libgst_cpu.so.  Prints one JSON line of results.
  tiny   : golden tiny cases (every `stride`-th), every D, one call each and
           all D variants in one batch; for the cases just before those, each
           D twice in a row (the pairs share a group-distance job); costs
           against the fixture, trees validated
  jobs   : one batch of random k=5 queries (duplicates allowed) over several
           group-distance jobs, against the oracle and against group_pass=0
  graph  : engine steps replayed from a captured graph against eager launches
  branching: a vertex that holds states of two or more groups only through
           grows, where the optimum needs a merge
  hop    : one batch of random k=3..5 queries with hop bounds 2..4 (depth-aware
           bounds and hop star / path upper bounds), then the same queries
           ordered by bound (runs of one bound share hop group-distance jobs),
           against the oracle
"""
import json
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from gpugst_amd import Graph, data  # noqa: E402
from gpugst_amd.validate import validate_tree  # noqa: E402


def groups_csr(groups):
    gptr = np.zeros(len(groups) + 1, np.int64)
    for i, m in enumerate(groups):
        gptr[i + 1] = gptr[i] + len(m)
    gmem = np.concatenate([np.asarray(m, np.int32) for m in groups])
    return gptr, gmem


def check_tiny(stride):
    cases = json.load(open(os.path.join(ROOT, "tests", "golden", "tiny.json")))
    bad, n = [], 0
    for ci in range(len(cases)):
        full, shared = ci % stride == 0, ci % stride == stride - 1
        if not (full or shared):
            continue
        c = cases[ci]
        rp, col, wt = data.csr_from_edges(c["V"], c["edges"])
        g = Graph(rp, col, wt)
        g.set_groups(*groups_csr(c["groups"]))
        ids = list(range(len(c["groups"])))
        todo = [(-1, c["cost"])] + [(int(D), hc) for D, hc in c["hop"].items()]
        if full:
            for D, exp in todo:
                r = g.solve(ids, D=D)
                n += 1
                ok = r.cost == exp
                if ok and r.feasible:
                    ok = validate_tree(rp, col, wt, c["groups"], r.cost, r.root, r.edges, D=D)["ok"]
                if not ok:
                    bad.append([ci, D, r.cost, exp])
            rs = g.solve_batch([ids] * len(todo), [d for d, _ in todo])
            for (D, exp), r in zip(todo, rs):
                n += 1
                if r.cost != exp:
                    bad.append([ci, D, r.cost, exp, "batch"])
        if shared:
            # each bound twice in a row: the pairs share a group-distance job,
            # and on these small graphs a hop job's frontier empties before its
            # last height
            rs = g.solve_batch([ids] * (2 * len(todo)), [d for d, _ in todo for _ in (0, 1)])
            for i, r in enumerate(rs):
                D, exp = todo[i // 2]
                n += 1
                if r.cost != exp:
                    bad.append([ci, D, r.cost, exp, "job"])
        g.close()
    return {"checked": n, "bad": bad}


def random_instance(V, seed, deg=3, gmax=4):
    rng = random.Random(seed)
    edges = set()
    for v in range(1, V):
        edges.add((rng.randrange(v), v))
    while len(edges) < deg * V:
        a, b = rng.randrange(V), rng.randrange(V)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    E = [(a, b, rng.randint(1, 20)) for a, b in sorted(edges)]
    rp, col, wt = data.csr_from_edges(V, E)
    G = 40
    groups = [sorted(rng.sample(range(V), rng.randint(1, gmax))) for _ in range(G)]
    return rng, rp, col, wt, groups


def check_jobs(V, nq, seed):
    rng, rp, col, wt, groups = random_instance(V, seed)
    G = len(groups)
    qs = [[rng.randrange(G) for _ in range(5)] for _ in range(nq)]
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    rs = g.solve_batch(qs)
    ctr = {x: g.get_option("batch_job_" + x) for x in ("entries", "arcs", "relaxations", "src_chunks")}
    ctr["hub_bytes"] = g.get_option("hub_bytes")
    g.set_option("hub16", 0)
    ctr["hub_bytes_32"] = g.get_option("hub_bytes")
    g.set_option("hub16", 1)
    g.set_option("backbone", 0)  # the same batch on every arc
    rf = g.solve_batch(qs)
    g.set_option("backbone", 1)
    g.set_option("group_pass", 0)
    r0 = g.solve_batch(qs)
    bad = []
    for i, q in enumerate(qs):
        gl = [groups[x] for x in q]
        exp = O.dpbf(rp, col, wt, gl)[0]
        ok = rs[i].cost == exp and r0[i].cost == exp and rf[i].cost == exp
        if ok and rs[i].feasible:
            ok = validate_tree(rp, col, wt, gl, rs[i].cost, rs[i].root, rs[i].edges)["ok"]
        if not ok:
            bad.append([i, rs[i].cost, r0[i].cost, exp])
    return {"checked": nq, "bad": bad, "labels": 5 * nq, "counters": ctr}


def check_hop(V, nq, seed):
    rng, rp, col, wt, groups = random_instance(V, seed, deg=4, gmax=8)
    G = len(groups)
    qs = [rng.sample(range(G), rng.randint(3, 5)) for _ in range(nq)]
    Ds = [rng.randint(2, 4) for _ in range(nq)]
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    rs = g.solve_batch(qs, Ds)
    # the same queries ordered by bound: runs of one D share hop jobs
    order = sorted(range(nq), key=lambda i: Ds[i])
    ro = g.solve_batch([qs[i] for i in order], [Ds[i] for i in order])
    rj = [None] * nq
    for pos, i in enumerate(order):
        rj[i] = ro[pos]
    bad, feas = [], 0
    for i, q in enumerate(qs):
        gl = [groups[x] for x in q]
        exp = O.dpbf(rp, col, wt, gl, D=Ds[i])[0]
        ok = rs[i].cost == exp and rj[i].cost == exp
        for r in (rs[i], rj[i]):
            if ok and r.feasible:
                ok = validate_tree(rp, col, wt, gl, r.cost, r.root, r.edges, D=Ds[i])["ok"]
        feas += int(ok and rs[i].feasible)
        if not ok:
            bad.append([i, Ds[i], rs[i].cost, rj[i].cost, exp])
    g.close()
    return {"checked": nq, "feasible": feas, "bad": bad}


def check_branching():
    """A vertex x where three two-group branches meet, holding its level-2
    states only through grows.  Each group also has a member in a second
    component, joined to one hub by edges of weight 7 (one of 6), so the star
    bound is 41 from the start; that bound prunes the single-group merges at x
    (22 + a rest bound of 20), while the branches grown into x pass (12 + 20).
    The optimum, 36, needs the four-group merge at x: without x among the
    merge vertices the DP answers the hub's star, 41."""
    E = []
    for b in range(3):
        E += [(0, 1 + b, 10), (1 + b, 4 + 2 * b, 1), (1 + b, 5 + 2 * b, 1)]
    E += [(10, 11 + i, 7 if i < 5 else 6) for i in range(6)]
    rp, col, wt = data.csr_from_edges(17, E)
    groups = [[4 + i, 11 + i] for i in range(6)]
    exp = O.dpbf(rp, col, wt, groups)[0]
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    r = g.solve(list(range(6)))
    ok = r.cost == exp and validate_tree(rp, col, wt, groups, r.cost, r.root, r.edges)["ok"]
    g.close()
    return {"cost": r.cost, "expect": exp, "ok": bool(ok)}


def check_backbone():
    """The graph's shortest-path backbone against the exact one (arcs with
    w == dist(u, v), from all-pairs distances): equal while every vertex is a
    hub (V <= 64), a superset beyond that; an arc and its reverse together.
    On two of the graphs, queries without a hop bound (one call each, batched,
    and batched on every arc) and one with D = 3 equal the oracle."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    out = {"graphs": 0, "bad": [], "removed": 0, "queries": 0}
    for gi, (V, seed) in enumerate([(40, 1), (64, 2), (150, 3), (220, 4)]):
        rng, rp, col, wt, groups = random_instance(V, 500 + seed, deg=4, gmax=3)
        dist = dijkstra(csr_matrix((wt.astype(np.float64), col, rp), shape=(V, V)))
        src = np.repeat(np.arange(V), np.diff(rp))
        exact = wt <= dist[src, col]
        g = Graph(rp, col, wt)
        g.set_groups(*groups_csr(groups))
        brp, bcol, bwt = g.backbone()
        bsrc = np.repeat(np.arange(V), np.diff(brp))
        kept = set(zip(bsrc.tolist(), bcol.tolist(), bwt.tolist()))
        want = set(zip(src[exact].tolist(), col[exact].tolist(), wt[exact].tolist()))
        allarcs = set(zip(src.tolist(), col.tolist(), wt.tolist()))
        out["graphs"] += 1
        out["removed"] += len(col) - len(bcol)
        if V <= 64 and kept != want:
            out["bad"].append([gi, "not exact", len(kept), len(want)])
        if not want <= kept or not kept <= allarcs:
            out["bad"].append([gi, "not a superset"])
        if any((v, u, w) not in kept for (u, v, w) in kept):
            out["bad"].append([gi, "asymmetric"])
        if g.get_option("backbone_arcs") != len(bcol):
            out["bad"].append([gi, "count"])
        if V not in (64, 220):  # queries on two of the graphs (the GPU tests take many more)
            g.close()
            continue
        G = len(groups)
        qs = [[rng.randrange(G) for _ in range(rng.randint(2, 5))] for _ in range(5)]
        Ds = [-1] * 4 + [3]
        g.set_option("backbone", 0)
        full = g.solve_batch(qs, Ds)
        g.set_option("backbone", 1)
        got = list(zip([g.solve(q, D=d) for q, d in zip(qs, Ds)], g.solve_batch(qs, Ds), full))
        for i, (q, d) in enumerate(zip(qs, Ds)):
            gl = [groups[x] for x in q]
            exp = O.dpbf(rp, col, wt, gl, D=d)[0]
            ok = all(r.cost == exp for r in got[i])
            for r in got[i]:
                if ok and r.feasible:
                    ok = validate_tree(rp, col, wt, gl, r.cost, r.root, r.edges, D=d)["ok"]
            out["queries"] += 1
            if not ok:
                out["bad"].append([gi, i, d, [r.cost for r in got[i]], exp])
        g.close()
    return out


def check_graph():
    """Batches with engine steps replayed from a captured graph (graph=1, the
    mock runtime records the launches) against eager launches and the oracle."""
    inst = data.make_config(2, scale=0.002, nq=2)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    qs = list(inst.queries)
    Ds = [3 if i % 4 == 1 else -1 for i in range(len(qs))]
    g.set_option("graph", 0)
    eager = [r.cost for r in g.solve_batch(qs, Ds)]
    g.set_option("graph", 1)
    rep = [[r.cost for r in g.solve_batch(qs, Ds)] for _ in range(2)]
    g.close()
    exp = [O.dpbf(inst.rowptr, inst.col, inst.wt, inst.query_groups(i), D=Ds[i])[0] for i in range(len(qs))]
    return {"expect": exp, "eager": eager, "graph": rep}


def check_static():
    """gst_solve_static (state in the library's static scratch, no device
    allocation) against gst_solve and the oracle: queries on a 150-vertex
    graph, then on a smaller graph (the same scratch, rebound), then the first
    graph again; an edge buffer that is too small is an error."""
    from gpugst_amd import GstError
    rng, rp, col, wt, groups = random_instance(150, 17, deg=3, gmax=4)
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    c = json.load(open(os.path.join(ROOT, "tests", "golden", "tiny.json")))[5]
    trp, tcol, twt = data.csr_from_edges(c["V"], c["edges"])
    t = Graph(trp, tcol, twt)
    t.set_groups(*groups_csr(c["groups"]))
    bad, n = [], 0
    plan = [(g, rp, col, wt, groups, [rng.randrange(40) for _ in range(4)], -1),
            (g, rp, col, wt, groups, [rng.randrange(40) for _ in range(3)], 3),
            (t, trp, tcol, twt, c["groups"], list(range(len(c["groups"]))), -1),
            (t, trp, tcol, twt, c["groups"], list(range(len(c["groups"]))), 2),
            (g, rp, col, wt, groups, [rng.randrange(40) for _ in range(5)], -1)]
    for gr, a, b, w, gs, q, D in plan:
        r = gr.solve_static(q, D=D)
        p = gr.solve(q, D=D)
        gl = [gs[x] for x in q]
        exp = O.dpbf(a, b, w, gl, D=D)[0]
        ok = r.cost == p.cost == exp
        if ok and r.feasible:
            ok = validate_tree(a, b, w, gl, r.cost, r.root, r.edges, D=D)["ok"]
        n += 1
        if not ok:
            bad.append([n, D, r.cost, p.cost, exp])
    small_err = False
    q = plan[0][5]
    need = g.solve(q).edges.shape[0]
    if need > 1:
        try:
            g.solve_static(q, edge_cap=need - 1)
        except GstError as e:
            small_err = e.status == 1
    g.close()
    t.close()
    return {"checked": n, "bad": bad, "small_buffer_error": small_err or need <= 1}


def check_robust():
    """Failure handling on the CPU build, forced from the host:
      * count_states fills states_finite (every singleton of a connected graph
        is finite) and max_frontier; -1 without the option;
      * a deadline stops a query on the device (status CANCELLED) and the
        engine answers the next query exactly;
      * a wait that times out (the mock stream reports busy) poisons the
        engine: the call and the next one fail with DEVICE_ERROR while the
        stream is busy, and once it drains the engine answers exactly;
      * the server: gst_cancel of a request, and gst_server_destroy with a
        thread blocked in gst_wait (it returns; nothing hangs)."""
    import ctypes
    import threading
    from gpugst_amd import GstError, Server
    from gpugst_amd.capi import GST_CANCELLED, GST_DEVICE_ERROR, lib_path
    rng, rp, col, wt, groups = random_instance(150, 23)
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    q = rng.sample(range(len(groups)), 5)
    exp = O.dpbf(rp, col, wt, [groups[x] for x in q])[0]
    out = {}
    r0 = g.solve(q)
    g.set_option("count_states", 1)
    r1 = g.solve(q)
    g.set_option("count_states", 0)
    out["count"] = [r0.counters["states_finite"], r1.counters["states_finite"], r1.counters["max_frontier"],
                    r1.cost == exp]
    out["count_ok"] = (r0.counters["states_finite"] == -1 and r1.counters["states_finite"] >= 150 * 5
                       and r1.counters["max_frontier"] > 0 and r0.cost == r1.cost == exp)
    # deadline
    g.set_option("deadline_ms", 1)
    st = None
    try:
        g.solve(q)
    except GstError as e:
        st = e.status
    rb = g.solve_batch([q, q[:3], q[1:]])
    g.set_option("deadline_ms", 0)
    out["deadline_ok"] = (st == GST_CANCELLED and all(r.status == GST_CANCELLED for r in rb)
                          and g.solve(q).cost == exp)
    # timeout -> poisoned until the stream drains
    mock = ctypes.CDLL(lib_path())
    g.set_option("timeout_ms", 30)
    mock.gsim_stream_busy(1)
    errs = []
    for _ in range(2):
        try:
            g.solve(q)
            errs.append(0)
        except GstError as e:
            errs.append(e.status)
    mock.gsim_stream_busy(0)
    g.set_option("timeout_ms", 60000)
    out["poison"] = errs
    out["poison_ok"] = errs == [GST_DEVICE_ERROR, GST_DEVICE_ERROR] and g.solve(q).cost == exp
    # server: a mix of bounds through the request loop equals one call each
    mq = [rng.sample(range(len(groups)), rng.randint(3, 5)) for _ in range(6)]
    mD = [-1, 2, -1, -1, 3, -1]
    srv = Server(g)
    ts = [srv.submit(x, D=d) for x, d in zip(mq, mD)]
    mixed = [srv.wait(t).cost for t in ts]
    out["server_stats"] = [srv.stat("steps"), srv.stat("chunks")]
    # group lists replaced while the server has requests in flight: the
    # change waits for the server's queued chunks, and nothing hangs
    t1 = [srv.submit(q) for _ in range(4)]
    g.set_groups(*groups_csr(groups))
    t2 = [srv.submit(q) for _ in range(4)]
    out["regroup_server"] = [srv.wait(t).cost for t in t1 + t2]
    out["regroup_server_ok"] = out["regroup_server"] == [exp] * 8
    srv.close()
    out["mix"] = mixed
    out["mix_ok"] = mixed == [O.dpbf(rp, col, wt, [groups[y] for y in x], D=d)[0] for x, d in zip(mq, mD)]
    # server cancel and destroy
    srv = Server(g)
    t = srv.submit(q)
    srv.cancel(t)
    rc = srv.wait(t)
    cancel_ok = rc.status == GST_CANCELLED or rc.cost == exp
    tickets = [srv.submit(q) for _ in range(12)]
    got = []
    th = threading.Thread(target=lambda: got.append(srv.wait(tickets[-1])))
    th.start()
    srv.close()
    th.join(timeout=120)
    out["server_ok"] = bool(cancel_ok and not th.is_alive() and len(got) == 1
                            and (got[0].status == GST_CANCELLED or got[0].cost == exp))
    g.close()
    return out


def check_cache():
    """Group-distance cache across batches (unbounded jobs): a group gets a
    column at its second request and is read from it afterwards (hits,
    labels counted per query); single-query jobs do the same; new group lists
    retire the cache; every cost equals the oracle."""
    rng, rp, col, wt, groups = random_instance(150, 29)
    G = len(groups)
    b1 = [rng.sample(range(G), 4) for _ in range(8)]
    b2 = [rng.sample(range(G), 4) for _ in range(8)]
    exp1 = [O.dpbf(rp, col, wt, [groups[x] for x in q])[0] for q in b1]
    exp2 = [O.dpbf(rp, col, wt, [groups[x] for x in q])[0] for q in b2]
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    g.set_option("cache_mb", 1)
    r1 = g.solve_batch(b1)   # first requests: no columns yet
    h1 = g.get_option("batch_cache_hits")
    i1 = g.get_option("batch_cache_inserts")
    r2 = g.solve_batch(b2)   # second requests of shared groups: columns written
    r3 = g.solve_batch(b1)   # third: read back
    h3 = g.get_option("batch_cache_hits") - h1
    out = {"hits_first": h1, "inserts_first": i1, "hits_later": h3, "inserts": g.get_option("batch_cache_inserts"),
           "per_query": sum(r.counters["cached_labels"] for r in r2 + r3)}
    out["exact"] = ([r.cost for r in r1] == exp1 and [r.cost for r in r2] == exp2
                    and [r.cost for r in r3] == exp1)
    # one query per batch: a job of one query fills the cache for the next
    g2 = Graph(rp, col, wt)
    g2.set_groups(*groups_csr(groups))
    g2.set_option("cache_mb", 1)
    g2.set_option("job_min_queries", 1)
    r4 = [g2.solve_batch([q])[0].cost for q in b1[:2] * 3]
    out["single_exact"] = r4 == exp1[:2] * 3 and g2.get_option("batch_cache_hits") >= 8
    # new group lists: the cache must not serve the old distances
    g.set_option("cache_mb", 1)
    g.solve_batch(b1)
    perm = groups[1:] + groups[:1]
    g.set_groups(*groups_csr(perm))
    r5 = g.solve_batch(b1)
    out["regroup_exact"] = [r.cost for r in r5] == [O.dpbf(rp, col, wt, [perm[x] for x in q])[0] for q in b1]
    g.close()
    g2.close()
    return out


def check_progressive():
    """Progressive answers (approx_num / approx_den): every answer is a valid
    tree (walk-trees allowed) whose cost is at least the optimum and at most
    num/den times the certified lower bound it reports, which is at most the
    optimum; ratio 1/1 gives the optimum; plain and batched, with and
    without a hop bound."""
    rng, rp, col, wt, groups = random_instance(150, 37, deg=3, gmax=4)
    G = len(groups)
    g = Graph(rp, col, wt)
    g.set_groups(*groups_csr(groups))
    qs = [rng.sample(range(G), rng.randint(3, 5)) for _ in range(4)]
    Ds = [-1, 3, -1, 2]
    opt = [O.dpbf(rp, col, wt, [groups[x] for x in q], D=d)[0] for q, d in zip(qs, Ds)]
    bad, early = [], 0
    for num, den in ((1, 1), (2, 1)):
        g.set_option("approx_num", num)
        g.set_option("approx_den", den)
        rs = [g.solve(q, D=d) for q, d in zip(qs, Ds)] + g.solve_batch(qs, Ds)
        for i, r in enumerate(rs):
            q, d, o = qs[i % 4], Ds[i % 4], opt[i % 4]
            lb = r.counters["lower_bound"]
            ok = r.cost >= o and lb <= o and r.cost * den <= num * lb
            if num == 1:
                ok = ok and r.cost == o
            if ok and r.feasible:
                ok = validate_tree(rp, col, wt, [groups[x] for x in q], r.cost, r.root, r.edges, D=d)["ok"]
            early += int(r.cost != lb)
            if not ok:
                bad.append([num, den, i, d, r.cost, lb, o])
    g.close()
    return {"bad": bad, "early": early}


def main():
    stride = int(sys.argv[1]) if len(sys.argv) > 1 else 3
    nq = int(sys.argv[2]) if len(sys.argv) > 2 else 60
    nh = int(sys.argv[3]) if len(sys.argv) > 3 else 24
    checks = {"tiny": lambda: check_tiny(stride), "jobs": lambda: check_jobs(150, nq, 7),
              "hop": lambda: check_hop(150, nh, 11), "branching": check_branching, "graph": check_graph,
              "static": check_static, "robust": check_robust, "cache": check_cache,
              "progressive": check_progressive, "backbone": check_backbone}
    # an optional fourth argument names the checks to run (comma-separated)
    names = sys.argv[4].split(",") if len(sys.argv) > 4 else list(checks)
    out = {name: checks[name]() for name in names}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
