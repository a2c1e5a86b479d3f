"""GPU parity: the HIP engine against the oracle and the golden fixtures.

Bar: costs bit-exact (integer DP); every returned tree validated (edges exist
with their weights, edge sum == cost, every query group covered, hop diameter
<= D).  Ties may pick different trees.
"""
import json
import os
import random

import numpy as np
import pytest

import oracle as O
from gpugst_amd import INFEASIBLE, Graph, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _graph_from(rp, col, wt, groups):
    g = Graph(rp, col, wt)
    gptr = np.zeros(len(groups) + 1, np.int64)
    for i, m in enumerate(groups):
        gptr[i + 1] = gptr[i] + len(m)
    gmem = np.concatenate([np.asarray(m, np.int32) for m in groups]) if groups else np.zeros(0, np.int32)
    g.set_groups(gptr, gmem)
    return g


def _check(rp, col, wt, groups, res, expect, D=-1):
    assert res.status == 0
    assert res.cost == expect, (res.cost, expect)
    if expect == INFEASIBLE:
        assert res.root == -1 and res.edges.shape[0] == 0
        return
    v = validate_tree(rp, col, wt, groups, res.cost, res.root, res.edges, D=D)
    assert v["ok"], v


def test_tiny_golden(golden_dir):
    cases = json.load(open(os.path.join(golden_dir, "tiny.json")))
    for ci, c in enumerate(cases):
        _say("tiny case", ci, "V", c["V"], "k", len(c["groups"]))
        rp, col, wt = data.csr_from_edges(c["V"], c["edges"])
        g = _graph_from(rp, col, wt, c["groups"])
        ids = list(range(len(c["groups"])))
        _check(rp, col, wt, c["groups"], g.solve(ids), c["cost"])
        for D, hc in c["hop"].items():
            _check(rp, col, wt, c["groups"], g.solve(ids, D=int(D)), hc, D=int(D))
        # each bound twice in a row: the pairs share a group-distance job, and on
        # these small graphs a hop job's frontier empties before its last height
        todo = [(-1, c["cost"])] + [(int(D), hc) for D, hc in c["hop"].items()]
        g.set_option("budget_mb", 16)  # the batch engine's state arena (default: most of the HBM)
        rs = g.solve_batch([ids] * (2 * len(todo)), [d for d, _ in todo for _ in (0, 1)])
        for i, r in enumerate(rs):
            D, hc = todo[i // 2]
            _check(rp, col, wt, c["groups"], r, hc, D=D)
        g.close()


def test_reference_instance(golden_dir):
    z = np.load(os.path.join(golden_dir, "simple_iterative_tests.npz"))
    gl = [z["gmem"][z["gptr"][i]:z["gptr"][i + 1]] for i in range(len(z["gptr"]) - 1)]
    g = Graph(z["rowptr"], z["col"], z["wt"])
    g.set_groups(z["gptr"], z["gmem"])
    for D, cost in zip(z["D"], z["cost"]):
        _say("reference instance D", int(D))
        _check(z["rowptr"], z["col"], z["wt"], gl, g.solve([0, 1, 2, 3], D=int(D)), int(cost), D=int(D))
    g.close()


def test_c1_all_queries(golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "c1_costs.json")))
    inst = data.make_config(1)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    for qi, q in enumerate(inst.queries):
        _say("C1 query", qi, list(q))
        r = g.solve(q)
        _check(inst.rowptr, inst.col, inst.wt, inst.query_groups(qi), r, ref["costs"][qi])
    g.close()


@pytest.mark.parametrize("seed", range(3))
def test_random_medium_vs_oracle(seed):
    rng = random.Random(1000 + seed)
    for it in range(6):
        V = rng.randint(50, 3000)
        m = rng.randint(V, 6 * V)
        u, v = data.chung_lu_edges(V, m, 2.3, seed * 100 + it)
        rp, col, wt, _ = data.weighted_csr(V, u, v, 1, rng.choice([4, 100, 1000]), seed * 100 + it + 7)
        G = 40
        gptr, gmem = data.make_groups(V, G, rng.randint(1, 8), 50, seed * 100 + it + 9)
        g = Graph(rp, col, wt)
        g.set_groups(gptr, gmem)
        for qi in range(5):
            k = rng.randint(1, 6)
            q = rng.sample(range(G), k)
            groups = [gmem[gptr[x]:gptr[x + 1]] for x in q]
            for D in (-1, rng.randint(0, 5)):
                _say("medium", seed, it, "V", V, "k", k, "D", D)
                expect = O.dpbf(rp, col, wt, groups, D=D)[0]
                _check(rp, col, wt, groups, g.solve(q, D=D), expect, D=D)
        g.close()


def test_infeasible_and_edge_cases():
    # two components; empty group; vertex in all groups; k = 1
    el = [(0, 1, 3), (1, 2, 4), (3, 4, 5)]
    rp, col, wt = data.csr_from_edges(6, el)
    groups = [[0], [2], [4], [], [1, 3], [5]]
    g = _graph_from(rp, col, wt, groups)
    _check(rp, col, wt, [[0], [2]], g.solve([0, 1]), 7)
    _check(rp, col, wt, [[0], [4]], g.solve([0, 2]), INFEASIBLE)
    _check(rp, col, wt, [[0], []], g.solve([0, 3]), INFEASIBLE)
    _check(rp, col, wt, [[1, 3]], g.solve([4]), 0)
    _check(rp, col, wt, [[5]], g.solve([5]), 0)          # isolated vertex, k = 1
    _check(rp, col, wt, [[5], [0]], g.solve([5, 0]), INFEASIBLE)
    _check(rp, col, wt, [[0], [2]], g.solve([0, 1], D=1), INFEASIBLE)
    _check(rp, col, wt, [[0], [2]], g.solve([0, 1], D=2), 7, D=2)
    _check(rp, col, wt, [[0], [0]], g.solve([0, 0]), 0)  # repeated group id
    g.close()


def test_vertex_reached_only_by_grows():
    # x = 0 joins three two-group branches; the star bound through the second
    # component (41) prunes the single-group merges at x, so x holds states of
    # two or more groups only through grows, and the optimum 36 needs the
    # four-group merge there (the same instance as the CPU-build check)
    E = []
    for b in range(3):
        E += [(0, 1 + b, 10), (1 + b, 4 + 2 * b, 1), (1 + b, 5 + 2 * b, 1)]
    E += [(10, 11 + i, 7 if i < 5 else 6) for i in range(6)]
    rp, col, wt = data.csr_from_edges(17, E)
    groups = [[4 + i, 11 + i] for i in range(6)]
    assert O.dpbf(rp, col, wt, groups)[0] == 36
    g = _graph_from(rp, col, wt, groups)
    _check(rp, col, wt, groups, g.solve(list(range(6))), 36)
    _check(rp, col, wt, groups, g.solve_batch([list(range(6))] * 3)[1], 36)
    g.close()


@pytest.mark.parametrize("D", [-1, 2, 3])
def test_batch_jobs_vs_oracle(D):
    # one batch of 16 queries of one bound: they share a group-distance job
    # (for D >= 0 a hop job with one row per height); every cost against the
    # oracle, every tree validated
    rng = random.Random(77 + D)
    V = 2500
    u, v = data.chung_lu_edges(V, 4 * V, 2.2, 900 + D)
    rp, col, wt, _ = data.weighted_csr(V, u, v, 1, 50, 901 + D)
    G = 60
    gptr, gmem = data.make_groups(V, G, 4, 40, 902 + D)
    g = Graph(rp, col, wt)
    g.set_groups(gptr, gmem)
    qs = [rng.sample(range(G), rng.randint(3, 5)) for _ in range(16)]
    _say("batch of 16, D", D)
    rs = g.solve_batch(qs, [D] * len(qs))
    assert g.get_option("batch_jobs") >= 1
    for q, r in zip(qs, rs):
        groups = [gmem[gptr[x]:gptr[x + 1]] for x in q]
        _check(rp, col, wt, groups, r, O.dpbf(rp, col, wt, groups, D=D)[0], D=D)
    g.close()


def test_batch_matches_plain():
    inst = data.make_config(2, scale=0.02, nq=6)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    qs = inst.queries
    Ds = [-1 if i % 3 else 2 for i in range(len(qs))]
    plain = [g.solve(q, D=d) for q, d in zip(qs, Ds)]
    _say("batch of", len(qs))
    batch = g.solve_batch(qs, Ds)
    for i, (a, b) in enumerate(zip(plain, batch)):
        groups = inst.query_groups(i)
        expect = O.dpbf(inst.rowptr, inst.col, inst.wt, groups, D=Ds[i])[0]
        _check(inst.rowptr, inst.col, inst.wt, groups, a, expect, D=Ds[i])
        _check(inst.rowptr, inst.col, inst.wt, groups, b, expect, D=Ds[i])
    g.close()


@pytest.mark.parametrize("D", [7, 15])
def test_k6_large_hop_bounds(D):
    # k = 6 with a large bound: 8 state heights (D = 15), the merge staging
    # holds few vertices per pass, and the merge predecessor uses its wide
    # height fields; plain and batched against the oracle
    rng = random.Random(600 + D)
    V = 120
    u, v = data.chung_lu_edges(V, 3 * V, 2.3, 610 + D)
    rp, col, wt, _ = data.weighted_csr(V, u, v, 1, 20, 611 + D)
    G = 30
    gptr, gmem = data.make_groups(V, G, 3, 10, 612 + D)
    g = Graph(rp, col, wt)
    g.set_groups(gptr, gmem)
    qs = [rng.sample(range(G), 6) for _ in range(4)]
    _say("k = 6, D", D)
    batch = g.solve_batch(qs, [D] * len(qs))
    for q, b in zip(qs, batch):
        groups = [gmem[gptr[x]:gptr[x + 1]] for x in q]
        expect = O.dpbf(rp, col, wt, groups, D=D)[0]
        _check(rp, col, wt, groups, g.solve(q, D=D), expect, D=D)
        _check(rp, col, wt, groups, b, expect, D=D)
    g.close()
