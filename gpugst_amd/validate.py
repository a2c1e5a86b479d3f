"""Validation of returned trees.

The reference's feasibility check only tests "connected" and "covers every
query group" (this_is_a_feasible_solution_gpu, code/PrunedDP++/include/
exp_CPU_nonHop.h:17-49).  This one also checks that every edge exists with its
weight, that the edge-weight sum equals the reported cost, the hop diameter
against D, and whether the witness is a proper tree (distinct vertices) or a
tree-like walk (possible only for the hop-bounded DP, whose optimum is taken
over walks).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np


def _has_arc(rowptr, col, wt, u, v, w) -> bool:
    lo, hi = int(rowptr[u]), int(rowptr[u + 1])
    seg = col[lo:hi]
    i = int(np.searchsorted(seg, v, side="left"))
    while i < seg.shape[0] and seg[i] == v:
        if wt[lo + i] == w:
            return True
        i += 1
    # rows may be unsorted in caller-provided graphs
    m = (seg == v) & (wt[lo:hi] == w)
    return bool(m.any())


def validate_tree(rowptr, col, wt, groups: Sequence[Sequence[int]], cost: int, root: int,
                  edges: np.ndarray, D: int = -1) -> dict:
    """Returns {'ok': bool, 'reason': str, 'walk': bool, 'diameter': int, 'sum': int}."""
    out = {"ok": False, "reason": "", "walk": False, "diameter": 0, "sum": 0}
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 4)
    n = edges.shape[0]
    if root < 0:
        out["reason"] = "no root"
        return out
    # parent structure: edge i hangs from the root (parent -1, u == root) or
    # from the child end of an earlier edge
    depth = np.zeros(n, dtype=np.int64)
    for i in range(n):
        u, v, w, par = (int(x) for x in edges[i])
        if par == -1:
            if u != root:
                out["reason"] = f"edge {i} claims the root but starts at {u} != {root}"
                return out
            depth[i] = 1
        else:
            if not (0 <= par < i) or int(edges[par, 1]) != u:
                out["reason"] = f"edge {i} has an inconsistent parent {par}"
                return out
            depth[i] = depth[par] + 1
        if not _has_arc(rowptr, col, wt, u, v, w):
            out["reason"] = f"edge ({u},{v},w={w}) is not in the graph"
            return out
    s = int(edges[:, 2].sum()) if n else 0
    out["sum"] = s
    if s != cost:
        out["reason"] = f"edge sum {s} != reported cost {cost}"
        return out
    verts = np.concatenate([[root], edges[:, 1]]) if n else np.asarray([root])
    vset = set(int(x) for x in verts)
    for gi, g in enumerate(groups):
        if not vset.intersection(int(x) for x in g):
            out["reason"] = f"query group {gi} is not covered"
            return out
    out["walk"] = len(vset) != verts.shape[0]
    # hop diameter of the (copy) tree: heights bottom-up
    height = np.zeros(n + 1, dtype=np.int64)  # index n = root copy
    best = 0
    top1 = np.zeros(n + 1, dtype=np.int64)
    top2 = np.zeros(n + 1, dtype=np.int64)
    for i in range(n - 1, -1, -1):
        h = height[i] + 1
        par = int(edges[i, 3])
        p = n if par == -1 else par
        if h > top1[p]:
            top2[p] = top1[p]
            top1[p] = h
        elif h > top2[p]:
            top2[p] = h
        height[p] = max(height[p], h)
    for i in range(n + 1):
        best = max(best, int(top1[i] + top2[i]))
    out["diameter"] = best
    if D is not None and D >= 0 and best > D:
        out["reason"] = f"hop diameter {best} > D = {D}"
        return out
    out["ok"] = True
    return out
