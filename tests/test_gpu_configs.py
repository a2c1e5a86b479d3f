"""GPU parity on two benchmark configurations at their stated sizes.

  - the 170,000-vertex Chung-Lu configuration (6.8M undirected edges, 13.6M
    arcs, exponent 2.0, weights in [1, 1000], 100,000 groups): 8 queries at
    k = 4 and 8 at k = 5, one call each and as one batch; every cost equals
    the oracle (exact pruned DPBF), every tree is valid;
  - the 3.0M-vertex serving configuration (117M undirected edges, 234M arcs,
    2M groups; k uniform in {3..6}, 20% of requests with D = 3): 16 requests
    submitted to the server at once equal one call per request, and every
    tree is valid with its bound.  An exact DP over 3M vertices is too slow
    for a test; the size-independent properties (mix == alone, valid trees
    with edge sums equal to the costs) carry the parity here, and the oracle
    pins the same engine on the smaller configurations.
"""
import time

import pytest

import oracle as O
from gpugst_amd import Graph, Server, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_config2_full_size_vs_oracle():
    t = time.time()
    inst = data.make_config(2, nq=8)
    assert inst.V == 170_000 and inst.col.shape[0] == 13_600_000
    _say("config 2: generated in", round(time.time() - t, 1), "s;", len(inst.queries), "queries")
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    qs = list(inst.queries)
    plain = []
    for i, q in enumerate(qs):
        _say("config 2: query", i, "k", len(q))
        plain.append(g.solve(q))
    batch = g.solve_batch(qs)
    for i, q in enumerate(qs):
        groups = inst.query_groups(i)
        expect = O.dpbf_pruned(inst.rowptr, inst.col, inst.wt, groups)[0]
        _say("config 2: oracle", i, expect)
        for r in (plain[i], batch[i]):
            assert r.status == 0 and r.cost == expect, (i, r.cost, expect)
            assert validate_tree(inst.rowptr, inst.col, inst.wt, groups, r.cost, r.root, r.edges)["ok"], i
    g.close()


def test_config5_full_size_mix_equals_alone():
    t = time.time()
    inst = data.make_config(5, nq=16)
    assert inst.V == 3_000_000 and inst.col.shape[0] == 234_000_000
    _say("config 5: generated in", round(time.time() - t, 1), "s")
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    qs, Ds = list(inst.queries), list(inst.D)
    assert any(d >= 0 for d in Ds) and any(d < 0 for d in Ds)
    srv = Server(g)
    tickets = [srv.submit(q, D=d) for q, d in zip(qs, Ds)]
    mix = [srv.wait(tk) for tk in tickets]
    srv.close()
    _say("config 5: mix done")
    for i, (q, d) in enumerate(zip(qs, Ds)):
        alone = g.solve(q, D=d)
        _say("config 5: request", i, "k", len(q), "D", d, "cost", alone.cost)
        assert mix[i].status == 0 and alone.status == 0
        assert mix[i].cost == alone.cost, (i, mix[i].cost, alone.cost)
        if alone.feasible:
            groups = inst.query_groups(i)
            for r in (mix[i], alone):
                v = validate_tree(inst.rowptr, inst.col, inst.wt, groups, r.cost, r.root, r.edges, D=d)
                assert v["ok"], (i, v)
    g.close()
