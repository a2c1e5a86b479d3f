"""Group-distance cache on the MI355X: columns of exact group distances kept
across batches and server epochs (option cache_mb).

The reference recomputes every query's group distances
(code/GPUGST/src/DPQ.cu:378-404, code/GPUGST+/src/GSTnonHop.cu:273); here a
group gets a column at its second request, and later batches or server
requests copy that column instead of recomputing it.  A 1 MiB budget holds fewer columns than the batches
use, so columns are evicted and re-inserted along the way.  Every cost must
equal the oracle's; the cache must be hit, and the per-query count of cached
labels must add up to the engine's hit counter.
"""
import random

import pytest

import oracle as O
from gpugst_amd import Graph, Server, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def test_cache_across_batches_and_server():
    inst = data.make_config(2, scale=0.02, nq=1)
    rng = random.Random(41)
    elig = [int(x) for x in inst.meta["eligible"][:90]]
    A = [rng.sample(elig, rng.randint(3, 5)) for _ in range(40)]
    B = [rng.sample(elig, rng.randint(3, 5)) for _ in range(40)]
    exp = {}
    for q in A + B:
        if tuple(q) not in exp:
            exp[tuple(q)] = O.dpbf(inst.rowptr, inst.col, inst.wt, [inst.group(x) for x in q])[0]
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    g.set_option("cache_mb", 1)
    for name, qs in (("A", A), ("B", B), ("A again", A)):
        h0 = g.get_option("batch_cache_hits")
        rs = g.solve_batch(qs)
        hits = g.get_option("batch_cache_hits") - h0
        print("cache: batch", name, "hits", hits, "evictions", g.get_option("batch_cache_evictions"),
              "entries", g.get_option("batch_cache_entries"), flush=True)
        assert sum(r.counters["cached_labels"] for r in rs) == hits
        for q, r in zip(qs, rs):
            assert r.cost == exp[tuple(q)], (q, r.cost, exp[tuple(q)])
            groups = [inst.group(x) for x in q]
            assert validate_tree(inst.rowptr, inst.col, inst.wt, groups, r.cost, r.root, r.edges)["ok"]
        if name == "A again":
            assert hits > 0
    assert g.get_option("batch_cache_evictions") > 0
    # the server engine keeps its own cache across epochs (every admission
    # through a job, so that lone requests fill it too)
    g.set_option("job_min_queries", 1)
    srv = Server(g)
    rounds = [[srv.wait(t) for t in [srv.submit(q) for q in B[:8]]] for _ in range(3)]
    srv.close()
    again = rounds[2]
    for rs in rounds:
        for q, r in zip(B[:8], rs):
            assert r.cost == exp[tuple(q)]
    print("cache: server cached labels", sum(r.counters["cached_labels"] for r in again), flush=True)
    assert sum(r.counters["cached_labels"] for r in again) > 0
    g.close()
