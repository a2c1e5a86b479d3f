"""The oracle is pinned before it is trusted (CPU only).

Each check is chosen so that a plausible restatement mistake fails it:
brute force over vertex subsets and Dreyfus-Wagner on the GST -> Steiner
reduction are independent algorithms for the unbounded optimum; closed forms
(k = 1, k = 2 as a multi-source shortest path, a vertex in every group, trees)
and hop invariants (monotone in D, D = 0, D >= V - 1 equals unbounded, k = 2 as a
hop-bounded shortest walk) bound the hop-bounded DP of CPUHOP.h; the center
formulation and brute force over proper subtrees in test_oracle_hop_pin.py pin
it exactly.
"""
import json
import os
import random

import numpy as np
import pytest

import oracle as O
from gpugst_amd import data
from gpugst_amd.validate import validate_tree


def _rand_graph(V, E, wmax, rng):
    edges = {}
    for v in range(1, V):
        u = rng.randrange(v)
        edges[(u, v)] = rng.randint(1, wmax)
    while len(edges) < E:
        u, v = rng.randrange(V), rng.randrange(V)
        if u != v:
            edges.setdefault((min(u, v), max(u, v)), rng.randint(1, wmax))
    return [(u, v, w) for (u, v), w in sorted(edges.items())]


def test_tiny_golden_brute_force_and_dw(golden_dir):
    cases = json.load(open(os.path.join(golden_dir, "tiny.json")))
    assert len(cases) >= 50
    for c in cases:
        rp, col, wt = data.csr_from_edges(c["V"], c["edges"])
        cost, root, edges, _ = O.dpbf(rp, col, wt, c["groups"])
        assert cost == c["cost"]
        assert O.brute_force_gst(c["V"], c["edges"], c["groups"]) == c["cost"]
        v = validate_tree(rp, col, wt, c["groups"], cost, root, np.array(edges).reshape(-1, 4))
        assert v["ok"] and not v["walk"], v
        for D, hc in c["hop"].items():
            assert O.dpbf(rp, col, wt, c["groups"], D=int(D))[0] == hc


@pytest.mark.parametrize("seed", range(4))
def test_random_against_dreyfus_wagner(seed):
    rng = random.Random(100 + seed)
    for _ in range(15):
        V = rng.randint(20, 120)
        el = _rand_graph(V, rng.randint(V - 1, 3 * V), 50, rng)
        k = rng.randint(2, 5)
        groups = [rng.sample(range(V), rng.randint(1, 4)) for _ in range(k)]
        rp, col, wt = data.csr_from_edges(V, el)
        cost, root, edges, _ = O.dpbf(rp, col, wt, groups)
        assert cost == O.gst_via_steiner(rp, col, wt, groups)
        assert validate_tree(rp, col, wt, groups, cost, root, np.array(edges).reshape(-1, 4))["ok"]


def test_closed_forms():
    rng = random.Random(5)
    for _ in range(20):
        V = rng.randint(10, 80)
        el = _rand_graph(V, rng.randint(V - 1, 2 * V), 30, rng)
        rp, col, wt = data.csr_from_edges(V, el)
        g1 = rng.sample(range(V), 3)
        g2 = rng.sample(range(V), 2)
        # k = 1: a member alone
        assert O.dpbf(rp, col, wt, [g1])[0] == 0
        # k = 2: distance between the groups (multi-source Dijkstra)
        d = O.sssp(rp, col, wt, g1)
        assert O.dpbf(rp, col, wt, [g1, g2])[0] == int(d[g2].min())
        # a vertex in every group costs nothing
        x = rng.randrange(V)
        assert O.dpbf(rp, col, wt, [g1 + [x], g2 + [x], [x]])[0] == 0
        # order of the query groups does not change the optimum
        g3 = rng.sample(range(V), 2)
        a = O.dpbf(rp, col, wt, [g1, g2, g3])[0]
        assert a == O.dpbf(rp, col, wt, [g3, g1, g2])[0]
        # lower and upper bounds: max group-pair distance <= OPT <= star
        ds = [O.sssp(rp, col, wt, g) for g in (g1, g2, g3)]
        assert max(int(ds[i][g].min()) for i, g in [(0, g2), (0, g3), (1, g3)]) <= a
        assert a <= int(np.min(ds[0] + ds[1] + ds[2]))


def test_tree_graph_unique_subtree():
    # on a tree the optimum is the minimal subtree spanning one member per group;
    # with singleton groups it is the Steiner subtree: sum of edges on paths
    rng = random.Random(9)
    for _ in range(20):
        V = rng.randint(5, 60)
        el = [(rng.randrange(v), v, rng.randint(1, 20)) for v in range(1, V)]
        rp, col, wt = data.csr_from_edges(V, el)
        terms = rng.sample(range(V), rng.randint(2, 5))
        parent = {v: (u, w) for u, v, w in el}
        depth = {0: 0}
        for u, v, w in el:
            depth[v] = depth[u] + 1
        used = set()
        for t in terms[1:]:
            a, b = t, terms[0]
            while a != b:
                if depth[a] < depth[b]:
                    a, b = b, a
                used.add(a)
                a = parent[a][0]
        # edges on the union of paths to terms[0] is the Steiner subtree only if
        # terms[0] lies on it; compute the true subtree by pruning leaves
        adj = {v: set() for v in range(V)}
        for u, v, w in el:
            adj[u].add(v)
            adj[v].add(u)
        keep = set(range(V))
        changed = True
        while changed:
            changed = False
            for v in list(keep):
                if v not in terms and len(adj[v] & keep) <= 1:
                    keep.discard(v)
                    changed = True
        expect = sum(w for u, v, w in el if u in keep and v in keep)
        assert O.dpbf(rp, col, wt, [[t] for t in terms])[0] == expect


def test_hop_invariants():
    rng = random.Random(11)
    for _ in range(25):
        V = rng.randint(6, 40)
        el = _rand_graph(V, rng.randint(V - 1, 2 * V), 20, rng)
        rp, col, wt = data.csr_from_edges(V, el)
        k = rng.randint(1, 4)
        groups = [rng.sample(range(V), rng.randint(1, 3)) for _ in range(k)]
        unb = O.dpbf(rp, col, wt, groups)[0]
        prev = O.INF
        for D in range(0, 8):
            c = O.dpbf(rp, col, wt, groups, D=D)[0]
            assert c <= prev          # non-increasing in D
            assert c >= unb           # a walk-tree never beats the tree optimum
            prev = c
        assert O.dpbf(rp, col, wt, groups, D=V - 1)[0] == unb
        # D = 0: zero iff one vertex lies in every group
        common = set(groups[0]).intersection(*map(set, groups[1:]))
        assert O.dpbf(rp, col, wt, groups, D=0)[0] == (0 if common else O.INF)
        # k = 2: cheapest walk of at most D hops between the groups (Bellman-Ford)
        if k >= 2:
            g1, g2 = groups[0], groups[1]
            for D in range(0, 5):
                dist = np.full(V, O.INF, np.int64)
                dist[g1] = 0
                best = int(dist[g2].min())
                for _h in range(D):
                    nd = dist.copy()
                    for u, v, w in el:
                        nd[v] = min(nd[v], dist[u] + w)
                        nd[u] = min(nd[u], dist[v] + w)
                    dist = nd
                    best = min(best, int(dist[g2].min()))
                got = O.dpbf(rp, col, wt, [g1, g2], D=D)[0]
                assert got == min(best, O.INF)


def test_hop_trees_validate():
    rng = random.Random(13)
    walks = proper = 0
    for _ in range(30):
        V = rng.randint(8, 40)
        el = _rand_graph(V, rng.randint(V - 1, 3 * V), 20, rng)
        rp, col, wt = data.csr_from_edges(V, el)
        groups = [rng.sample(range(V), rng.randint(1, 3)) for _ in range(rng.randint(2, 4))]
        for D in (1, 2, 3, 4):
            c, root, edges, _ = O.dpbf(rp, col, wt, groups, D=D)
            if c == O.INF:
                continue
            v = validate_tree(rp, col, wt, groups, c, root, np.array(edges).reshape(-1, 4), D=D)
            assert v["ok"], v
            walks += v["walk"]
            proper += not v["walk"]
    # witnesses are valid either way; most are proper trees (the equality of
    # those with brute force is checked in test_oracle_hop_pin.py)
    assert proper > walks


def test_reference_instance_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "simple_iterative_tests.npz"))
    gl = [z["gmem"][z["gptr"][i]:z["gptr"][i + 1]] for i in range(len(z["gptr"]) - 1)]
    assert len(gl) == 4 and z["rowptr"].shape[0] == 10001 and z["col"].shape[0] == 100000
    for D, cost in zip(z["D"], z["cost"]):
        assert O.dpbf(z["rowptr"], z["col"], z["wt"], gl, D=int(D))[0] == int(cost)
    assert O.gst_via_steiner(z["rowptr"], z["col"], z["wt"], gl) == int(z["cost"][0])


def test_c1_fixture(golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "c1_costs.json")))
    inst = data.make_config(1)
    from tests.golden.make_golden import instance_digest
    assert instance_digest(inst) == ref["digest"], "generator drift"
    for qi in range(0, len(inst.queries), 7):
        assert O.dpbf(inst.rowptr, inst.col, inst.wt, inst.query_groups(qi))[0] == ref["costs"][qi]


def test_planted_mistake_is_caught():
    """A restatement error planted in the oracle's output (a dropped edge, a
    wrong cost) must fail the validator."""
    inst = data.make_config(1)
    gs = inst.query_groups(0)
    c, root, edges, _ = O.dpbf(inst.rowptr, inst.col, inst.wt, gs)
    e = np.array(edges).reshape(-1, 4)
    assert validate_tree(inst.rowptr, inst.col, inst.wt, gs, c, root, e)["ok"]
    assert not validate_tree(inst.rowptr, inst.col, inst.wt, gs, c + 1, root, e)["ok"]
    assert not validate_tree(inst.rowptr, inst.col, inst.wt, gs, c, root, e[:-1])["ok"]
    bad = e.copy()
    bad[0, 2] += 1
    assert not validate_tree(inst.rowptr, inst.col, inst.wt, gs, c + 1, root, bad)["ok"]


def test_pruned_cpu_solver_matches_unpruned():
    """The pruned best-first solver (the CPU baseline) must give the same
    optimum as the unpruned oracle."""
    rng = random.Random(21)
    for _ in range(60):
        V = rng.randint(3, 80)
        u, v = data.chung_lu_edges(V, min(rng.randint(V, 3 * V), V * (V - 1) // 2), 2.3, rng.randrange(1 << 30))
        rp, col, wt, _ = data.weighted_csr(V, u, v, 1, rng.choice([3, 20, 100]), rng.randrange(1 << 30))
        groups = [rng.sample(range(V), rng.randint(1, 3)) for _ in range(rng.randint(1, 6))]
        assert O.dpbf(rp, col, wt, groups)[0] == O.dpbf_pruned(rp, col, wt, groups)[0]
    inst = data.make_config(1)
    for qi in range(0, 50, 5):
        gs = inst.query_groups(qi)
        assert O.dpbf(inst.rowptr, inst.col, inst.wt, gs)[0] == O.dpbf_pruned(inst.rowptr, inst.col, inst.wt, gs)[0]
