"""CPU oracle for the group Steiner tree DP -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
library (``gpugst_amd``) never calls it.

The C restatement lives in ``gst_oracle.c`` (reference file:line citations in
its header); this module is a thin ctypes wrapper plus small pure-Python
brute-force checkers for tiny graphs.
"""
from __future__ import annotations

import ctypes
import itertools
import os
import subprocess
from typing import Sequence

import numpy as np

INF = 0x3FFFFFFF
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgst_oracle.so")
_lib = None


class OrcEdge(ctypes.Structure):
    _fields_ = [("u", ctypes.c_int32), ("v", ctypes.c_int32),
                ("w", ctypes.c_int32), ("parent", ctypes.c_int32)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.orc_dpbf.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int32, P, P,
                               ctypes.c_int32, P, P, P, ctypes.c_int32, P, P]
        L.orc_dpbf.restype = ctypes.c_int
        L.orc_sssp.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int64, P, P]
        L.orc_sssp.restype = ctypes.c_int
        L.orc_steiner_dw.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int32, P, P]
        L.orc_steiner_dw.restype = ctypes.c_int
        L.orc_dpbf_pruned.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int32, P, P, P, P]
        L.orc_dpbf_pruned.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _csr(rowptr, col, wt):
    return (np.ascontiguousarray(rowptr, dtype=np.int64),
            np.ascontiguousarray(col, dtype=np.int32),
            np.ascontiguousarray(wt, dtype=np.int32))


def load_variant(path: str):
    """A ctypes handle on another build of gst_oracle.c (the tests compile
    copies with a planted mistake to show that their checks catch it)."""
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.orc_dpbf.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int32, P, P,
                           ctypes.c_int32, P, P, P, ctypes.c_int32, P, P]
    L.orc_dpbf.restype = ctypes.c_int
    return L


def dpbf(rowptr, col, wt, groups: Sequence[Sequence[int]], D: int = -1,
         edge_cap: int | None = None, library=None):
    """Exact optimum of the GST DP.  Returns (cost, root, edges, stats).

    ``groups`` are the member lists of the k query groups (bit i = groups[i]).
    ``D < 0`` is the unbounded DP, ``D >= 0`` the hop-bounded one.  ``edges`` is
    a list of (parent, child, weight, parent_edge_index) tuples.  cost == INF
    means infeasible.  ``library`` selects another build (load_variant).
    """
    rowptr, col, wt = _csr(rowptr, col, wt)
    V = rowptr.shape[0] - 1
    k = len(groups)
    gptr = np.zeros(k + 1, dtype=np.int64)
    for i, g in enumerate(groups):
        gptr[i + 1] = gptr[i] + len(g)
    gmem = np.ascontiguousarray(
        np.concatenate([np.asarray(g, dtype=np.int32) for g in groups])
        if k else np.zeros(0, np.int32), dtype=np.int32)
    cap = edge_cap if edge_cap is not None else max(16, 4 * V)
    edges = (OrcEdge * cap)()
    cost = ctypes.c_int32()
    root = ctypes.c_int32()
    ne = ctypes.c_int32()
    stats = np.zeros(2, dtype=np.int64)
    rc = (library or lib()).orc_dpbf(V, _ptr(rowptr), _ptr(col), _ptr(wt), k, _ptr(gptr),
                        _ptr(gmem), D, ctypes.byref(cost), ctypes.byref(root),
                        edges, cap, ctypes.byref(ne), _ptr(stats))
    if rc < 0:
        raise RuntimeError(f"orc_dpbf failed rc={rc}")
    out = [(edges[i].u, edges[i].v, edges[i].w, edges[i].parent)
           for i in range(min(ne.value, cap))]
    return cost.value, root.value, out, stats


def dpbf_pruned(rowptr, col, wt, groups: Sequence[Sequence[int]]):
    """Exact optimum (unbounded) by the pruned best-first search; cost only.
    stats: pops, heap capacity, precompute ns, main-loop ns."""
    rowptr, col, wt = _csr(rowptr, col, wt)
    V = rowptr.shape[0] - 1
    k = len(groups)
    gptr = np.zeros(k + 1, dtype=np.int64)
    for i, g in enumerate(groups):
        gptr[i + 1] = gptr[i] + len(g)
    gmem = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.int32) for g in groups]),
                                dtype=np.int32)
    cost = ctypes.c_int32()
    stats = np.zeros(4, dtype=np.int64)
    rc = lib().orc_dpbf_pruned(V, _ptr(rowptr), _ptr(col), _ptr(wt), k, _ptr(gptr), _ptr(gmem),
                               ctypes.byref(cost), _ptr(stats))
    if rc < 0:
        raise RuntimeError(f"orc_dpbf_pruned failed rc={rc}")
    return cost.value, stats


def sssp(rowptr, col, wt, sources) -> np.ndarray:
    rowptr, col, wt = _csr(rowptr, col, wt)
    V = rowptr.shape[0] - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    dist = np.empty(V, dtype=np.int32)
    rc = lib().orc_sssp(V, _ptr(rowptr), _ptr(col), _ptr(wt), src.shape[0],
                        _ptr(src), _ptr(dist))
    if rc:
        raise RuntimeError(f"orc_sssp failed rc={rc}")
    return dist


def steiner_dw(rowptr, col, wt, terminals) -> int:
    rowptr, col, wt = _csr(rowptr, col, wt)
    V = rowptr.shape[0] - 1
    term = np.ascontiguousarray(terminals, dtype=np.int32)
    out = ctypes.c_int64()
    rc = lib().orc_steiner_dw(V, _ptr(rowptr), _ptr(col), _ptr(wt),
                              term.shape[0], _ptr(term), ctypes.byref(out))
    if rc:
        raise RuntimeError(f"orc_steiner_dw failed rc={rc}")
    return int(out.value)


def gst_via_steiner(rowptr, col, wt, groups) -> int:
    """GST optimum through the reduction to Steiner tree: one dummy terminal per
    group joined to every member with weight M > sum of all weights; the
    Steiner optimum is then OPT + k*M.  Returns INF if infeasible."""
    rowptr, col, wt = _csr(rowptr, col, wt)
    V = rowptr.shape[0] - 1
    k = len(groups)
    M = int(wt.astype(np.int64).sum()) + 1
    if any(len(g) == 0 for g in groups):
        return INF
    if k == 1:
        return 0  # a single terminal is a zero-cost Steiner tree by itself
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(rowptr))
    us = [src, ]
    vs = [col.astype(np.int64)]
    ws = [wt.astype(np.int64)]
    for i, g in enumerate(groups):
        g = np.unique(np.asarray(g, dtype=np.int64))
        d = np.full(g.shape[0], V + i, dtype=np.int64)
        us += [d, g]
        vs += [g, d]
        ws += [np.full(g.shape[0], M, np.int64)] * 2
    u = np.concatenate(us)
    v = np.concatenate(vs)
    w = np.concatenate(ws)
    order = np.lexsort((v, u))
    u, v, w = u[order], v[order], w[order]
    n = V + k
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, u + 1, 1)
    rp = np.cumsum(rp)
    if int(w.max(initial=0)) * 4 * (k + 1) >= (1 << 30):
        raise ValueError("weights too large for the Dreyfus-Wagner cross-check")
    c = steiner_dw(rp, v.astype(np.int32), w.astype(np.int32),
                   np.arange(V, V + k, dtype=np.int32))
    if c < 0:
        return INF
    opt = c - k * M
    return opt if opt < M else INF


# ---------------------------------------------------------------------------
# brute force on tiny graphs (pure Python)
# ---------------------------------------------------------------------------

def _mst_cost(vset, adj):
    """Prim over the induced subgraph; None if disconnected."""
    vset = list(vset)
    if len(vset) == 1:
        return 0
    inside = {vset[0]}
    sset = set(vset)
    total = 0
    while len(inside) < len(sset):
        best = None
        for a in inside:
            for b, w in adj[a]:
                if b in sset and b not in inside and (best is None or w < best[0]):
                    best = (w, b)
        if best is None:
            return None
        total += best[0]
        inside.add(best[1])
    return total


def brute_force_gst(V, edges_uvw, groups):
    """Exact GST by enumerating vertex subsets (V <= ~14): the optimum tree is
    the MST of the induced subgraph of its vertex set."""
    adj = [[] for _ in range(V)]
    for u, v, w in edges_uvw:
        adj[u].append((v, w))
        adj[v].append((u, w))
    gsets = [set(g) for g in groups]
    best = INF
    for r in range(1, V + 1):
        for S in itertools.combinations(range(V), r):
            ss = set(S)
            if not all(gs & ss for gs in gsets):
                continue
            c = _mst_cost(S, adj)
            if c is not None and c < best:
                best = c
    return best


# ---------------------------------------------------------------------------
# hop-bounded optimum, two independent ways (pure Python, small graphs)
# ---------------------------------------------------------------------------

def _popcount(x: int) -> int:
    return bin(x).count("1")


def hop_center_dp(V, edges_uvw, groups, D: int) -> int:
    """Cheapest tree-like walk of hop diameter <= D covering every group,
    computed around its center -- a formulation independent of the rooted
    recurrence of CPUHOP.h:192-227 that the oracle restates.

    A tree of diameter <= 2r has a center vertex c with every vertex within r
    hops of c; a tree of diameter 2r + 1 has a center arc (a, b) whose two
    sides lie within r hops of a and of b.  F_h(v, p) is the cheapest tree
    rooted at v, of height <= h, covering the groups of p: height 0 is v alone
    (p within v's groups), height h hangs branches (arc v-u plus a tree of
    height <= h - 1 at u) and closes under disjoint unions at v.  The answer is
    min_v F_r(v, full) for D = 2r, and additionally min over arcs (a, b) and
    splits p1 | p2 = full of F_r(a, p1) + w + F_r(b, p2) for D = 2r + 1.
    Walk-trees may revisit vertices, as the DP's states do."""
    if D < 0:
        raise ValueError("D must be >= 0")
    k = len(groups)
    W = 1 << k
    full = W - 1
    gmask = [0] * V
    for i, g in enumerate(groups):
        for v in g:
            gmask[int(v)] |= 1 << i
    nbr = [[] for _ in range(V)]
    for u, v, w in edges_uvw:
        nbr[u].append((v, w))
        nbr[v].append((u, w))
    order = sorted(range(1, W), key=_popcount)
    F = [[0 if (p & ~gmask[v]) == 0 else INF for p in range(W)] for v in range(V)]
    for _h in range(1, D // 2 + 1):
        prev = F
        F = []
        for v in range(V):
            row = [INF] * W
            row[0] = 0
            for p in order:
                b = 0 if (p & ~gmask[v]) == 0 else prev[v][p]
                for u, w in nbr[v]:
                    b = min(b, prev[u][p] + w)
                low = p & -p
                s = (p - 1) & p
                while s:
                    if s & low:
                        b = min(b, row[s] + row[p ^ s])
                    s = (s - 1) & p
                row[p] = min(b, INF)
            F.append(row)
    best = min(F[v][full] for v in range(V))
    if D % 2 == 1:
        for u, v, w in edges_uvw:
            for p1 in range(1, full):
                best = min(best, F[u][p1] + w + F[v][full ^ p1])
    return best if best < INF else INF


def brute_force_hop_trees(V, edges_uvw, groups, Dmax: int):
    """For D = 0..Dmax, the cheapest proper tree (a set of graph edges, no
    vertex repeated) of hop diameter <= D covering every group, by enumerating
    every subtree of the graph (tiny graphs only: each subtree is reached once
    by deciding the boundary edges in order, include or exclude).  Returns a
    list indexed by D; INF where no such tree exists.  The DP's optimum ranges
    over tree-like walks, so it can only be <= these values, and equal when
    its witness is a proper tree."""
    k = len(groups)
    full = (1 << k) - 1
    gmask = [0] * V
    for i, g in enumerate(groups):
        for v in g:
            gmask[int(v)] |= 1 << i
    E = [(int(u), int(v), int(w)) for u, v, w in edges_uvw]
    inc = [[] for _ in range(V)]
    for i, (u, v, _w) in enumerate(E):
        inc[u].append(i)
        inc[v].append(i)
    best = [INF] * (Dmax + 1)
    if any(m == full for m in gmask):
        return [0] * (Dmax + 1)

    def report(verts, tedges):
        cov = 0
        for x in verts:
            cov |= gmask[x]
        if cov != full:
            return
        cost = sum(E[i][2] for i in tedges)
        adj = {x: [] for x in verts}
        for i in tedges:
            u, v, _w = E[i]
            adj[u].append(v)
            adj[v].append(u)
        diam = 0
        for s in verts:
            dist = {s: 0}
            frontier = [s]
            while frontier:
                nxt = []
                for x in frontier:
                    for y in adj[x]:
                        if y not in dist:
                            dist[y] = dist[x] + 1
                            nxt.append(y)
                frontier = nxt
            diam = max(diam, max(dist.values()))
        for D in range(diam, Dmax + 1):
            if cost < best[D]:
                best[D] = cost

    def rec(r, verts, tedges, cand):
        if not cand:
            report(verts, tedges)
            return
        e, rest = cand[0], cand[1:]
        rec(r, verts, tedges, rest)  # e excluded for good
        u, v, _w = E[e]
        x = v if u in verts else u
        nv = verts | {x}
        nc = [f for f in rest if not (E[f][0] in nv and E[f][1] in nv)]
        for f in inc[x]:
            y = E[f][0] if E[f][1] == x else E[f][1]
            if y > r and y not in nv:
                nc.append(f)
        rec(r, nv, tedges + [e], nc)

    for r in range(V):
        cand = []
        for f in inc[r]:
            y = E[f][0] if E[f][1] == r else E[f][1]
            if y > r:
                cand.append(f)
        rec(r, {r}, [], cand)
    return best
