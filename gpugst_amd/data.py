"""Inputs: seeded synthetic instances for the benchmark configs and readers /
writers of the reference's on-disk formats.

Formats (reference README.md:4-20):
* ``X.in``       -- first line "V E", then "u v w" per undirected edge
  (read_input_graph, code/GPUGST/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:75-148)
* ``X_beg_pos.bin`` / ``X_csr.bin`` / ``X_weight.bin`` -- int64 arrays of the
  CSR (code/GPUGST+/include/graph.hpp:12-118; V = filesize/8 - 1)
* ``X.g``        -- line "gN:m1 m2 ..." is group N-1 (read_Group, same header :190-239)
* ``XK.csv``     -- one query per line, K group ids separated by spaces
  (read_inquire, :240-286)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import build as _build

_dlib = None


def _data_lib():
    global _dlib
    if _dlib is None:
        if not os.path.exists(_build.LIBDATA):
            raise OSError(f"{_build.LIBDATA} is missing: run `python -m gpugst_amd.build`")
        L = ctypes.CDLL(_build.LIBDATA)
        P, I32, I64, U64, F64 = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                 ctypes.c_uint64, ctypes.c_double)
        L.gstd_chung_lu.argtypes = [I64, I64, F64, U64, P, P]
        L.gstd_chung_lu.restype = I64
        L.gstd_rmat.argtypes = [I64, I64, I32, F64, F64, F64, U64, P, P]
        L.gstd_rmat.restype = I64
        L.gstd_sort_edges.argtypes = [I64, P, P]
        L.gstd_sort_edges.restype = None
        L.gstd_weights.argtypes = [I64, I32, I32, U64, P]
        L.gstd_weights.restype = None
        L.gstd_csr.argtypes = [I64, I64, P, P, P, P, P, P]
        L.gstd_csr.restype = I64
        L.gstd_components.argtypes = [I64, P, P, P, P]
        L.gstd_components.restype = I64
        L.gstd_group_sizes.argtypes = [I64, I64, I64, U64, P]
        L.gstd_group_sizes.restype = None
        L.gstd_group_members.argtypes = [I64, I64, P, U64, P]
        L.gstd_group_members.restype = None
        L.gstd_queries.argtypes = [I64, I32, P, I64, U64, P]
        L.gstd_queries.restype = None
        L.gstd_uniform_ints.argtypes = [I64, I64, I64, U64, P]
        L.gstd_uniform_ints.restype = None
        L.gstd_exponential.argtypes = [I64, F64, U64, P]
        L.gstd_exponential.restype = None
        PP = ctypes.POINTER(ctypes.c_void_p)
        L.gstd_read_in.argtypes = [ctypes.c_char_p, P, P, PP, PP, PP, P]
        L.gstd_read_in.restype = ctypes.c_int
        L.gstd_read_groups.argtypes = [ctypes.c_char_p, P, PP, PP, P]
        L.gstd_read_groups.restype = ctypes.c_int
        L.gstd_read_queries.argtypes = [ctypes.c_char_p, P, PP, PP, P]
        L.gstd_read_queries.restype = ctypes.c_int
        L.gstd_free.argtypes = [P]
        L.gstd_free.restype = None
        _dlib = L
    return _dlib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Instance:
    rowptr: np.ndarray
    col: np.ndarray
    wt: np.ndarray
    gptr: np.ndarray
    gmem: np.ndarray
    queries: List[np.ndarray] = field(default_factory=list)
    D: List[int] = field(default_factory=list)
    meta: dict = field(default_factory=dict)

    @property
    def V(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    def group(self, g: int) -> np.ndarray:
        return self.gmem[self.gptr[g]:self.gptr[g + 1]]

    def query_groups(self, qi: int) -> List[np.ndarray]:
        return [self.group(int(g)) for g in self.queries[qi]]


# ---------------------------------------------------------------- graphs

def chung_lu_edges(V: int, m: int, gamma: float, seed: int):
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    n = _data_lib().gstd_chung_lu(V, m, gamma, seed, _p(u), _p(v))
    if n < 0:
        raise ValueError("bad Chung-Lu parameters")
    return u[:n], v[:n]


def rmat_edges(V: int, m: int, scale: int, a=0.57, b=0.19, c=0.19, seed: int = 0):
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    n = _data_lib().gstd_rmat(V, m, scale, a, b, c, seed, _p(u), _p(v))
    if n < 0:
        raise ValueError("bad R-MAT parameters")
    return u[:n], v[:n]


def weighted_csr(V: int, u, v, wlo: int, whi: int, wseed: int):
    """Canonical edge order, seeded uniform weights, CSR with both arcs."""
    u = np.ascontiguousarray(u, np.int32).copy()
    v = np.ascontiguousarray(v, np.int32).copy()
    m = u.shape[0]
    L = _data_lib()
    L.gstd_sort_edges(m, _p(u), _p(v))
    w = np.empty(m, np.int32)
    L.gstd_weights(m, wlo, whi, wseed, _p(w))
    rowptr = np.empty(V + 1, np.int64)
    col = np.empty(2 * m, np.int32)
    wt = np.empty(2 * m, np.int32)
    L.gstd_csr(V, m, _p(u), _p(v), _p(w), _p(rowptr), _p(col), _p(wt))
    return rowptr, col, wt, (u, v, w)


def components(rowptr, col):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    V = rowptr.shape[0] - 1
    comp = np.empty(V, np.int32)
    sz = ctypes.c_int64()
    lcc = _data_lib().gstd_components(V, _p(rowptr), _p(col), _p(comp), ctypes.byref(sz))
    return comp, int(lcc), int(sz.value)


def restrict_to_lcc(V, u, v, w, rowptr, col):
    comp, lcc, size = components(rowptr, col)
    keep = comp == lcc
    newid = -np.ones(V, np.int64)
    newid[keep] = np.arange(size)
    sel = keep[u] & keep[v]
    nu, nv, nw = newid[u[sel]].astype(np.int32), newid[v[sel]].astype(np.int32), w[sel]
    rp = np.empty(size + 1, np.int64)
    c = np.empty(2 * nu.shape[0], np.int32)
    wt = np.empty(2 * nu.shape[0], np.int32)
    _data_lib().gstd_csr(size, nu.shape[0], _p(nu), _p(nv), _p(np.ascontiguousarray(nw)),
                         _p(rp), _p(c), _p(wt))
    return size, rp, c, wt


# ---------------------------------------------------------------- groups

def make_groups(V: int, G: int, mean: int, maxsize: int, seed: int):
    L = _data_lib()
    sizes = np.empty(G, np.int64)
    L.gstd_group_sizes(G, mean, maxsize, seed, _p(sizes))
    gptr = np.zeros(G + 1, np.int64)
    np.cumsum(sizes, out=gptr[1:])
    gmem = np.empty(int(gptr[-1]), np.int32)
    L.gstd_group_members(V, G, _p(gptr), seed ^ 0xA5A5A5A5, _p(gmem))
    return gptr, gmem


def eligible_groups(gptr, gmem, in_lcc: np.ndarray) -> np.ndarray:
    hit = in_lcc[gmem].astype(np.int64)
    cs = np.concatenate([[0], np.cumsum(hit)])
    cnt = cs[gptr[1:]] - cs[gptr[:-1]]
    return np.nonzero(cnt > 0)[0].astype(np.int32)


def make_queries(nq: int, k: int, eligible: np.ndarray, seed: int) -> np.ndarray:
    out = np.empty((nq, k), np.int32)
    el = np.ascontiguousarray(eligible, np.int32)
    _data_lib().gstd_queries(nq, k, _p(el), el.shape[0], seed, _p(out))
    return out


def uniform_ints(n: int, lo: int, hi: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    _data_lib().gstd_uniform_ints(n, lo, hi, seed, _p(out))
    return out


def exponential(n: int, mean: float, seed: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _data_lib().gstd_exponential(n, mean, seed, _p(out))
    return out


# ---------------------------------------------------------------- configs

# The five workloads of the benchmark plan; seeds are 1000*c + {1 graph,
# 2 weights, 3 groups, 4 queries, 5 arrivals}.
CONFIGS = {
    1: dict(name="chunglu-20k-k3", gen="chung_lu", V=20_000, m=150_000, gamma=2.1, w=(1, 100),
            lcc=True, G=5_000, gmean=5, gmax=200, queries=[(3, 50)], D=-1),
    2: dict(name="chunglu-170k-k45", gen="chung_lu", V=170_000, m=6_800_000, gamma=2.0, w=(1, 1000),
            lcc=False, G=100_000, gmean=8, gmax=1_000, queries=[(4, 300), (5, 300)], D=-1),
    3: dict(name="rmat-1.1M-k5-D4", gen="rmat", V=1_100_000, m=3_000_000, scale=21, w=(1, 100),
            lcc=False, G=500_000, gmean=6, gmax=1_000, queries=[(5, 300)], D=4),
    4: dict(name="rmat-1.1M-kmix-batched", gen="rmat", V=1_100_000, m=3_000_000, scale=21, w=(1, 100),
            lcc=False, G=500_000, gmean=6, gmax=1_000, queries="mix", nq=4096, kmix=(3, 6), D=-1,
            graph_of=3),
    5: dict(name="chunglu-3M-serving", gen="chung_lu", V=3_000_000, m=117_000_000, gamma=2.0, w=(1, 1000),
            lcc=False, G=2_000_000, gmean=10, gmax=5_000, queries="mix", nq=2000, kmix=(3, 6), D=-1,
            hop_frac=0.2, hop_D=3),
}


def make_config(c: int, scale: float = 1.0, nq: Optional[int] = None) -> Instance:
    """Build config c.  ``scale`` < 1 shrinks V, m and G proportionally (the
    parity tests use scaled-down versions); ``nq`` overrides query counts."""
    cfg = dict(CONFIGS[c])
    V = max(16, int(cfg["V"] * scale))
    m = max(V, int(cfg["m"] * scale))
    G = max(8, int(cfg["G"] * scale))
    base = 1000 * c
    gbase = 1000 * cfg.get("graph_of", c)  # configs sharing one graph share its seeds
    if cfg["gen"] == "chung_lu":
        u, v = chung_lu_edges(V, m, cfg["gamma"], gbase + 1)
    else:
        sc = int(cfg["scale"])
        while (1 << (sc - 1)) >= V and sc > 1:
            sc -= 1
        u, v = rmat_edges(V, m, sc, seed=gbase + 1)
    rowptr, col, wt, (su, sv, sw) = weighted_csr(V, u, v, cfg["w"][0], cfg["w"][1], gbase + 2)
    if cfg["lcc"]:
        V, rowptr, col, wt = restrict_to_lcc(V, su, sv, sw, rowptr, col)
    gptr, gmem = make_groups(V, G, cfg["gmean"], cfg["gmax"], gbase + 3)
    comp, lcc, _ = components(rowptr, col)
    elig = eligible_groups(gptr, gmem, comp == lcc)
    inst = Instance(rowptr, col, wt, gptr, gmem,
                    meta=dict(config=c, name=cfg["name"], V=V, arcs=int(col.shape[0]),
                              groups=G, scale=scale))
    inst.meta["eligible"] = elig
    if cfg["queries"] == "mix":
        n = nq if nq is not None else cfg["nq"]
        ks = uniform_ints(n, cfg["kmix"][0], cfg["kmix"][1], base + 4)
        qs = []
        for i in range(n):
            qs.append(make_queries(1, int(ks[i]), elig, (base + 4) * 1_000_003 + i)[0])
        inst.queries = qs
        if "hop_frac" in cfg:
            sel = uniform_ints(n, 0, 999, base + 6) < int(cfg["hop_frac"] * 1000)
            inst.D = [cfg["hop_D"] if s else -1 for s in sel]
        else:
            inst.D = [cfg["D"]] * n
    else:
        for k, n in cfg["queries"]:
            n = nq if nq is not None else n
            inst.queries += list(make_queries(n, k, elig, base + 4 + 10 * k))
        inst.D = [cfg["D"]] * len(inst.queries)
    return inst


# ---------------------------------------------------------------- formats

def write_in(path: str, V: int, u, v, w):
    with open(path, "w") as f:
        f.write(f"{V} {len(u)}\n")
        for a, b, c in zip(u, v, w):
            f.write(f"{a} {b} {c}\n")


def _take_array(L, ptr, n, dtype):
    """Copies n elements of a library-owned array into numpy and frees it."""
    try:
        if n == 0 or not ptr.value:
            return np.zeros(0, dtype)
        ct = ctypes.c_int64 if dtype == np.int64 else ctypes.c_int32
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(n,)).astype(dtype, copy=True)
    finally:
        if ptr.value:
            L.gstd_free(ptr)


def _reader_error(rc, path, bad):
    if rc == -1:
        raise OSError(f"cannot read {path}")
    if rc == -2:
        raise ValueError(f"{path}: malformed line {bad}")
    raise MemoryError(f"{path}: out of memory")


def read_edges(path: str):
    """Edge-list text file -> (V, u, v, w) arrays, parsed by the native reader
    (multithreaded: a 117M-edge file loads in seconds)."""
    L = _data_lib()
    V, m, bad = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    pu, pv, pw = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    rc = L.gstd_read_in(path.encode(), ctypes.byref(V), ctypes.byref(m), ctypes.byref(pu), ctypes.byref(pv),
                        ctypes.byref(pw), ctypes.byref(bad))
    if rc != 0:
        _reader_error(rc, path, bad.value)
    n = m.value
    return (V.value, _take_array(L, pu, n, np.int32), _take_array(L, pv, n, np.int32),
            _take_array(L, pw, n, np.int32))


def read_in(path: str):
    """Edge-list text format -> CSR (both arcs, rows sorted by neighbour)."""
    V, u, v, w = read_edges(path)
    rowptr = np.empty(V + 1, np.int64)
    col = np.empty(2 * u.shape[0], np.int32)
    wt = np.empty(2 * u.shape[0], np.int32)
    _data_lib().gstd_csr(V, u.shape[0], _p(u), _p(v), _p(w), _p(rowptr), _p(col), _p(wt))
    return rowptr, col, wt


def write_bin(prefix: str, rowptr, col, wt):
    np.asarray(rowptr, np.int64).tofile(prefix + "_beg_pos.bin")
    np.asarray(col, np.int64).tofile(prefix + "_csr.bin")
    np.asarray(wt, np.int64).tofile(prefix + "_weight.bin")


def read_bin(prefix: str):
    """int64 beg_pos / csr / weight files -> int64 rowptr, int32 col / wt
    (graph.hpp:12-118: V = size of beg_pos / 8 - 1, arcs = beg_pos[V]; the
    32-bit narrowing there is checked here)."""
    rowptr = np.fromfile(prefix + "_beg_pos.bin", dtype=np.int64)
    if rowptr.shape[0] < 2 or rowptr[0] != 0 or np.any(np.diff(rowptr) < 0):
        raise ValueError(f"{prefix}_beg_pos.bin is not a row-pointer array")
    col = np.fromfile(prefix + "_csr.bin", dtype=np.int64) if os.path.exists(prefix + "_csr.bin") else None
    wt = np.fromfile(prefix + "_weight.bin", dtype=np.int64) if os.path.exists(prefix + "_weight.bin") else None
    V = rowptr.shape[0] - 1
    if col is not None:
        if col.shape[0] != rowptr[-1]:
            raise ValueError("csr length does not match beg_pos[V]")
        if col.shape[0] and (col.min() < 0 or col.max() >= V):
            raise ValueError("csr entry outside [0, V)")
        col = col.astype(np.int32)
    if wt is not None:
        if wt.shape[0] != rowptr[-1]:
            raise ValueError("weight length does not match beg_pos[V]")
        if wt.shape[0] and (wt.min() < -(1 << 31) or wt.max() >= (1 << 31)):
            raise ValueError("weight outside int32")
        wt = wt.astype(np.int32)
    return rowptr, col, wt


def write_groups(path: str, gptr, gmem):
    with open(path, "w") as f:
        for g in range(len(gptr) - 1):
            mem = " ".join(str(int(x)) for x in gmem[gptr[g]:gptr[g + 1]])
            f.write(f"g{g + 1}:{mem}\n")


def read_groups(path: str):
    """.g file -> (gptr, gmem); line "gN:..." is group N-1, members sorted
    (native multithreaded reader)."""
    L = _data_lib()
    G, bad = ctypes.c_int64(), ctypes.c_int64()
    pp, pm = ctypes.c_void_p(), ctypes.c_void_p()
    rc = L.gstd_read_groups(path.encode(), ctypes.byref(G), ctypes.byref(pp), ctypes.byref(pm), ctypes.byref(bad))
    if rc != 0:
        _reader_error(rc, path, bad.value)
    gptr = _take_array(L, pp, G.value + 1, np.int64)
    gmem = _take_array(L, pm, int(gptr[-1]) if gptr.shape[0] else 0, np.int32)
    if gptr.shape[0] == 0:
        gptr = np.zeros(1, np.int64)
    return gptr, gmem


def write_queries(path: str, queries):
    with open(path, "w") as f:
        for q in queries:
            f.write(" ".join(str(int(x)) for x in q) + "\n")


def read_queries(path: str) -> List[np.ndarray]:
    """Query file -> one int32 array of group ids per non-blank line."""
    L = _data_lib()
    nq, bad = ctypes.c_int64(), ctypes.c_int64()
    pp, pi = ctypes.c_void_p(), ctypes.c_void_p()
    rc = L.gstd_read_queries(path.encode(), ctypes.byref(nq), ctypes.byref(pp), ctypes.byref(pi), ctypes.byref(bad))
    if rc != 0:
        _reader_error(rc, path, bad.value)
    qptr = _take_array(L, pp, nq.value + 1, np.int64)
    ids = _take_array(L, pi, int(qptr[-1]) if qptr.shape[0] else 0, np.int32)
    return [ids[qptr[i]:qptr[i + 1]].copy() for i in range(nq.value)]


def read_gstp(path: str):
    """GSTP text instance (graph_hash_of_mixed_weighted_read_for_GSTP,
    code/GPUGST/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:9-73):
    'input_graph Vertex v w', 'input_graph Edge u v w', 'group_graph Edge g v 1',
    'group_vertices g'.  Returns (V, edges (u,v,w), {group vertex: members},
    query group vertices)."""
    verts, edges, gedges, gverts = [], [], [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) < 2:
                continue
            if p[0] == "input_graph" and p[1] == "Vertex":
                verts.append(int(p[2]))
            elif p[0] == "input_graph" and p[1] == "Edge":
                edges.append((int(p[2]), int(p[3]), int(float(p[4]))))
            elif p[0] == "group_graph" and p[1] == "Edge":
                gedges.append((int(p[2]), int(p[3])))
            elif p[0] == "group_vertices":
                gverts.append(int(p[1]))
    V = max(verts) + 1 if verts else 0
    gset = set(gverts)
    groups = {g: [] for g in gverts}
    for a, b in gedges:
        if a in gset and b not in gset:
            groups[a].append(b)
        elif b in gset and a not in gset:
            groups[b].append(a)
    return V, edges, {g: sorted(set(m)) for g, m in groups.items()}, gverts


def csr_from_edges(V: int, edges):
    u = np.asarray([e[0] for e in edges], np.int32)
    v = np.asarray([e[1] for e in edges], np.int32)
    w = np.asarray([e[2] for e in edges], np.int32)
    rowptr = np.empty(V + 1, np.int64)
    col = np.empty(2 * u.shape[0], np.int32)
    wt = np.empty(2 * u.shape[0], np.int32)
    _data_lib().gstd_csr(V, u.shape[0], _p(u), _p(v), _p(w), _p(rowptr), _p(col), _p(wt))
    return rowptr, col, wt
