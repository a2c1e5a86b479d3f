"""The shortest-path backbone claim, pinned on the oracle (CPU only).

The engine runs queries without a hop bound on the graph's shortest-path
backbone: the arcs (u, v, w) with no u-v path shorter than w.  The claim is
that no optimal group Steiner tree needs a removed arc (replacing one by the
shorter path and breaking the cycles this closes leaves a cheaper connected
subgraph over the same vertices), so the optimum on the backbone equals the
optimum on the graph.  Checked here with the oracle on both graphs, each
backbone computed exactly from all-pairs distances (scipy); and the
counter-example that keeps hop-bounded queries on every arc.
"""
import random

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

import oracle as O
from gpugst_amd import data


def _rand_edges(V, E, wmax, rng):
    edges = {}
    for v in range(1, V):
        edges[(rng.randrange(v), v)] = rng.randint(1, wmax)
    while len(edges) < E:
        u, v = rng.randrange(V), rng.randrange(V)
        if u != v:
            edges.setdefault((min(u, v), max(u, v)), rng.randint(1, wmax))
    return [(u, v, w) for (u, v), w in sorted(edges.items())]


def exact_backbone(V, edges):
    """Edges (u, v, w) with w == dist(u, v) (all-pairs Dijkstra)."""
    rp, col, wt = data.csr_from_edges(V, edges)
    dist = dijkstra(csr_matrix((wt.astype(np.float64), col, rp), shape=(V, V)))
    return [(u, v, w) for (u, v, w) in edges if w <= dist[u, v]]


def test_backbone_keeps_every_optimum():
    rng = random.Random(2024)
    removed_total = 0
    for _ in range(40):
        V = rng.randint(12, 60)
        edges = _rand_edges(V, rng.randint(V, 4 * V), 60, rng)
        bb = exact_backbone(V, edges)
        removed_total += len(edges) - len(bb)
        rp, col, wt = data.csr_from_edges(V, edges)
        brp, bcol, bwt = data.csr_from_edges(V, bb)
        # the backbone keeps every distance
        full_d = O.sssp(rp, col, wt, [0])
        assert np.array_equal(full_d, O.sssp(brp, bcol, bwt, [0]))
        for _ in range(4):
            k = rng.randint(2, 5)
            groups = [rng.sample(range(V), rng.randint(1, 3)) for _ in range(k)]
            assert O.dpbf(rp, col, wt, groups)[0] == O.dpbf(brp, bcol, bwt, groups)[0]
    assert removed_total > 200  # the graphs do have redundant arcs


def test_backbone_needs_all_arcs_under_a_hop_bound():
    """a-b of weight 10 is redundant (a-c-b costs 2) but is the only way to
    join {a} and {b} within one hop: hop-bounded queries keep every arc."""
    edges = [(0, 1, 10), (0, 2, 1), (1, 2, 1)]
    bb = exact_backbone(3, edges)
    assert (0, 1, 10) not in bb
    rp, col, wt = data.csr_from_edges(3, edges)
    brp, bcol, bwt = data.csr_from_edges(3, bb)
    groups = [[0], [1]]
    assert O.dpbf(rp, col, wt, groups)[0] == O.dpbf(brp, bcol, bwt, groups)[0] == 2
    assert O.dpbf(rp, col, wt, groups, D=1)[0] == 10
    assert O.dpbf(brp, bcol, bwt, groups, D=1)[0] == O.INF
