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


def dpbf(rowptr, col, wt, groups: Sequence[Sequence[int]], D: int = -1,
         edge_cap: int | None = None):
    """Exact optimum of the GST DP.  Returns (cost, root, edges, stats).

    ``groups`` are the member lists of the k query groups (bit i = groups[i]).
    ``D < 0`` is the unbounded DP, ``D >= 0`` the hop-bounded one.  ``edges`` is
    a list of (parent, child, weight, parent_edge_index) tuples.  cost == INF
    means infeasible.
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
    rc = lib().orc_dpbf(V, _ptr(rowptr), _ptr(col), _ptr(wt), k, _ptr(gptr),
                        _ptr(gmem), D, ctypes.byref(cost), ctypes.byref(root),
                        edges, cap, ctypes.byref(ne), _ptr(stats))
    if rc < 0:
        raise RuntimeError(f"orc_dpbf failed rc={rc}")
    out = [(edges[i].u, edges[i].v, edges[i].w, edges[i].parent)
           for i in range(min(ne.value, cap))]
    return cost.value, root.value, out, stats


def dpbf_pruned(rowptr, col, wt, groups: Sequence[Sequence[int]]):
    """Exact optimum (unbounded) by the pruned best-first search; cost only."""
    rowptr, col, wt = _csr(rowptr, col, wt)
    V = rowptr.shape[0] - 1
    k = len(groups)
    gptr = np.zeros(k + 1, dtype=np.int64)
    for i, g in enumerate(groups):
        gptr[i + 1] = gptr[i] + len(g)
    gmem = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.int32) for g in groups]),
                                dtype=np.int32)
    cost = ctypes.c_int32()
    stats = np.zeros(2, dtype=np.int64)
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
