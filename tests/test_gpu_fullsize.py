"""GPU parity at the benchmark sizes, through properties that do not depend on
size: the batched engine and the server return the plain engine's costs,
every tree is valid, and a bounded sample matches the oracle.

  - the 1.1M-vertex R-MAT graph (6.0M arcs, 500,000 groups) with 96 queries of
    mixed k in {3..6}, unbounded: one batch (several group-distance jobs)
    against one call per query; the oracle (exact pruned DPBF) on the first
    two queries with k <= 4
  - the same graph, k = 5, hop bound D = 4: one batch against one call per
    query (an exact hop DP over 1.1M vertices is too slow for a test; the hop
    oracle is checked on the smaller graphs of test_gpu_parity.py and below)
  - the serving configuration's shape (Chung-Lu, exponent 2.0, weights in
    [1, 1000], mixed k in {3..6}, 20% of requests with D = 3) scaled to 30,000
    vertices and 2.34M arcs: every request submitted to the server at once
    (mix), against one call per query (alone) and against the oracle
"""
import numpy as np
import pytest

import oracle as O
from gpugst_amd import Graph, Server, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _valid(inst, i, r, D=-1):
    if not r.feasible:
        return True
    return validate_tree(inst.rowptr, inst.col, inst.wt, inst.query_groups(i), r.cost, r.root, r.edges,
                         D=D)["ok"]


@pytest.fixture(scope="module")
def rmat():
    inst = data.make_config(4, nq=96)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    yield inst, g
    g.close()


def test_rmat_mixed_k_batch_equals_plain(rmat):
    inst, g = rmat
    qs = inst.queries
    _say("rmat mixed k: batch of", len(qs))
    batch = g.solve_batch(qs)
    for i, q in enumerate(qs):
        r = g.solve(q)
        assert r.status == 0 and batch[i].status == 0
        assert batch[i].cost == r.cost, (i, list(q), batch[i].cost, r.cost)
        assert _valid(inst, i, r) and _valid(inst, i, batch[i]), (i, list(q))
    checked = 0
    for i, q in enumerate(qs):
        if len(q) > 4 or checked == 2:
            continue
        _say("rmat mixed k: oracle on query", i, "k", len(q))
        assert O.dpbf_pruned(inst.rowptr, inst.col, inst.wt, inst.query_groups(i))[0] == batch[i].cost
        checked += 1
    assert checked == 2


def test_rmat_hop_batch_equals_plain(rmat):
    inst, g = rmat
    qs = data.make_queries(24, 5, inst.meta["eligible"], 3000 + 4 + 50)
    _say("rmat D=4: batch of", len(qs))
    batch = g.solve_batch(list(qs), [4] * len(qs))
    feas = 0
    for i, q in enumerate(qs):
        r = g.solve(q, D=4)
        assert batch[i].cost == r.cost, (i, list(q), batch[i].cost, r.cost)
        groups = [inst.group(int(x)) for x in q]
        for res in (r, batch[i]):
            if res.feasible:
                v = validate_tree(inst.rowptr, inst.col, inst.wt, groups, res.cost, res.root, res.edges, D=4)
                assert v["ok"], (i, v)
        feas += int(r.feasible)
    assert feas > 0


def test_serving_shape_mix_equals_alone():
    inst = data.make_config(5, scale=0.01, nq=40)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    srv = Server(g)
    _say("serving shape: V", inst.meta["V"], "arcs", inst.meta["arcs"], "requests", len(inst.queries))
    tickets = [srv.submit(q, D=d) for q, d in zip(inst.queries, inst.D)]
    mix = [srv.wait(t) for t in tickets]
    srv.close()
    assert any(d >= 0 for d in inst.D) and any(d < 0 for d in inst.D)
    for i, (q, d) in enumerate(zip(inst.queries, inst.D)):
        alone = g.solve(q, D=d)
        _say("serving shape: request", i, "k", len(q), "D", d, "cost", alone.cost)
        assert mix[i].cost == alone.cost, (i, mix[i].cost, alone.cost)
        assert O.dpbf(inst.rowptr, inst.col, inst.wt, inst.query_groups(i), D=d)[0] == alone.cost
        assert _valid(inst, i, mix[i], D=d) and _valid(inst, i, alone, D=d)
    g.close()
    _ = np
