"""GPU: the shortest-path backbone each graph gets at creation.

Queries without a hop bound run on the arcs (u, v, w) that no shorter u-v path
replaces; hub distances computed on the device decide which arcs go.  Checked
here at full size against exact single-source distances of the oracle: every
arc with w == dist(u, v) is kept and every arc left out has a shorter path.
Costs with the backbone on and off agree with each other and with the oracle,
and the 20K-vertex configuration runs on every arc against its golden costs.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from gpugst_amd import Graph, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def rmat():
    inst = data.make_config(3, nq=1)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    yield inst, g
    g.close()


def test_rmat_backbone_against_exact_distances(rmat):
    inst, g = rmat
    rp, col, wt = g.backbone()
    n = col.shape[0]
    _say("backbone arcs", n, "of", inst.col.shape[0], "rounds", g.get_option("backbone_rounds"))
    assert 0 < n < inst.col.shape[0]
    assert g.get_option("backbone_arcs") == n
    deg = np.diff(inst.rowptr)
    rng = np.random.default_rng(7)
    srcs = list(rng.choice(np.flatnonzero(deg >= 2), 5, replace=False)) + [int(np.argmax(deg))]
    for s in srcs:
        _say("source", s, "degree", int(deg[s]))
        d = O.sssp(inst.rowptr, inst.col, inst.wt, [s])
        a0, a1 = inst.rowptr[s], inst.rowptr[s + 1]
        full = list(zip(inst.col[a0:a1].tolist(), inst.wt[a0:a1].tolist()))
        kept = list(zip(col[rp[s]:rp[s + 1]].tolist(), wt[rp[s]:rp[s + 1]].tolist()))
        assert sorted(kept) == kept and set(kept) <= set(full)
        for (v, w) in full:
            if w == d[v]:
                assert (v, w) in kept, (s, v, w)
            elif (v, w) not in kept:
                assert d[v] < w, (s, v, w, d[v])
        # an arc and its reverse are kept together
        for (v, w) in kept[:64]:
            back = list(zip(col[rp[v]:rp[v + 1]].tolist(), wt[rp[v]:rp[v + 1]].tolist()))
            assert (s, w) in back


def test_rmat_costs_with_and_without_backbone(rmat):
    inst, g = rmat
    qs = data.make_queries(24, 5, inst.meta["eligible"], 91)
    g.set_option("backbone", 0)
    full = g.solve_batch(list(qs))
    g.set_option("backbone", 1)
    bb = g.solve_batch(list(qs))
    plain = [g.solve(q) for q in qs[:6]]
    assert [r.cost for r in full] == [r.cost for r in bb]
    assert [r.cost for r in plain] == [r.cost for r in bb[:6]]
    for i in range(3):
        groups = [inst.group(x) for x in qs[i]]
        exp = O.dpbf_pruned(inst.rowptr, inst.col, inst.wt, groups)[0]
        _say("query", i, "cost", bb[i].cost, "oracle", exp)
        assert bb[i].cost == exp
    for r, q in zip(bb + plain, list(qs) + list(qs[:6])):
        groups = [inst.group(x) for x in q]
        v = validate_tree(inst.rowptr, inst.col, inst.wt, groups, r.cost, r.root, r.edges)
        assert v["ok"], v


@pytest.mark.parametrize("nq", [28, 32, 38])
def test_rmat_hub_bound_fills_agree(rmat, nq):
    """Unbounded jobs start their rows at hub bounds, filled in packed 16-bit
    arithmetic (this graph's hub distances allow it), in 32-bit arithmetic,
    or not at all (rows at infinity): the same costs, equal to the oracle.
    28, 32 and 38 queries of k = 5 make one job of 144, 160 and 192 labels,
    whose last group of 16, 32 and 64 labels the 16-bit fill takes with 8,
    16 and 32 lanes per vertex."""
    inst, g = rmat
    qs = data.make_queries(nq, 5, inst.meta["eligible"], 93)
    costs = {}
    for hb, h16 in ((1, 1), (1, 0), (0, 0)):
        g.set_option("hub_bounds", hb)
        g.set_option("hub16", h16)
        assert g.get_option("hub_bytes") == (2 if h16 else 4)
        costs[(hb, h16)] = [r.cost for r in g.solve_batch(list(qs))]
        _say("hub_bounds", hb, "hub16", h16, "costs", costs[(hb, h16)][:6])
    g.set_option("hub_bounds", 1)
    g.set_option("hub16", 1)
    assert costs[(1, 1)] == costs[(1, 0)] == costs[(0, 0)]
    for i in range(2):
        groups = [inst.group(x) for x in qs[i]]
        assert costs[(1, 1)][i] == O.dpbf_pruned(inst.rowptr, inst.col, inst.wt, groups)[0]


def test_c1_every_arc_vs_golden(golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "c1_costs.json")))
    inst = data.make_config(1)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    g.set_option("backbone", 0)
    res = g.solve_batch(list(inst.queries))
    for qi, r in enumerate(res):
        assert r.cost == ref["costs"][qi], (qi, r.cost, ref["costs"][qi])
    for qi in range(5):
        r = g.solve(inst.queries[qi])
        assert r.cost == ref["costs"][qi]
        v = validate_tree(inst.rowptr, inst.col, inst.wt, inst.query_groups(qi), r.cost, r.root, r.edges)
        assert v["ok"], v
    g.close()
