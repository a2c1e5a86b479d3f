"""ctypes binding of ``include/gst.h`` (libgst.so).

This is a thin host-side mirror of the C ABI used by the tests, the benchmark
and Python callers.  It never falls back to anything: if the native library is
missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import build as _build

GST_OK = 0
GST_INVALID_ARG = 1
GST_OUT_OF_MEMORY = 2
GST_DEVICE_ERROR = 3
GST_CANCELLED = 4
GST_NOT_READY = 5
GST_QUEUE_FULL = 6
INFEASIBLE = 0x3FFFFFFF
UNBOUNDED = -1
MAX_K = 6
MAX_D = 15

STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "DEVICE_ERROR",
                4: "CANCELLED", 5: "NOT_READY", 6: "QUEUE_FULL"}

# every function include/gst.h declares
EXPORTED = ("gst_graph_create", "gst_graph_destroy", "gst_backbone_get", "gst_groups_set", "gst_solve",
            "gst_solve_static", "gst_solve_batch", "gst_result_free", "gst_set_option", "gst_get_option",
            "gst_query_bytes", "gst_server_create", "gst_submit", "gst_poll", "gst_wait",
            "gst_cancel", "gst_server_stat", "gst_server_destroy", "gst_last_error", "gst_version")


class GstError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class c_edge(ctypes.Structure):
    _fields_ = [("u", ctypes.c_int32), ("v", ctypes.c_int32),
                ("w", ctypes.c_int32), ("parent", ctypes.c_int32)]


class c_counters(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("frontier_entries", ctypes.c_int64),
                ("arcs", ctypes.c_int64), ("relaxations", ctypes.c_int64),
                ("improvements", ctypes.c_int64), ("merge_evals", ctypes.c_int64),
                ("live_vertices", ctypes.c_int64), ("shared_entries", ctypes.c_int64),
                ("shared_arcs", ctypes.c_int64), ("shared_relaxations", ctypes.c_int64),
                ("active_vertices", ctypes.c_int64), ("max_frontier", ctypes.c_int64),
                ("states_finite", ctypes.c_int64), ("cached_labels", ctypes.c_int64),
                ("lower_bound", ctypes.c_int64)]


class c_result(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_int32), ("root", ctypes.c_int32),
                ("n_edges", ctypes.c_int32), ("status", ctypes.c_int32),
                ("edges", ctypes.POINTER(c_edge)), ("counters", c_counters),
                ("latency_ms", ctypes.c_double), ("device_ms", ctypes.c_double)]


class c_query(ctypes.Structure):
    _fields_ = [("group_ids", ctypes.POINTER(ctypes.c_int32)), ("k", ctypes.c_int32),
                ("D", ctypes.c_int32)]


_lib = None


def lib_path() -> str:
    # GST_LIB selects another build of the same ABI (the CPU build used by tests)
    return os.environ.get("GST_LIB", _build.LIBGST)


def load():
    """Load libgst.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: run `python -m gpugst_amd.build` first")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I32, I64 = ctypes.c_int32, ctypes.c_int64
    sig = {
        "gst_graph_create": [P, I64, P, P, I64, ctypes.POINTER(P)],
        "gst_graph_destroy": [P],
        "gst_backbone_get": [P, P, P, P, ctypes.POINTER(I64)],
        "gst_groups_set": [P, P, P, I64],
        "gst_solve": [P, P, I32, I32, ctypes.POINTER(c_result)],
        "gst_solve_batch": [P, I64, P, P],
        "gst_solve_static": [P, P, I32, I32, P, I32, ctypes.POINTER(c_result)],
        "gst_set_option": [P, ctypes.c_char_p, I64],
        "gst_get_option": [P, ctypes.c_char_p, ctypes.POINTER(I64)],
        "gst_query_bytes": [P, I32, I32, ctypes.POINTER(I64)],
        "gst_server_create": [P, I64, I32, ctypes.POINTER(P)],
        "gst_submit": [P, ctypes.POINTER(c_query), ctypes.POINTER(ctypes.c_uint64)],
        "gst_poll": [P, ctypes.c_uint64, ctypes.POINTER(c_result)],
        "gst_wait": [P, ctypes.c_uint64, ctypes.POINTER(c_result)],
        "gst_cancel": [P, ctypes.c_uint64],
        "gst_server_destroy": [P],
        "gst_server_stat": [P, ctypes.c_char_p, ctypes.POINTER(I64)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.gst_result_free.argtypes = [ctypes.POINTER(c_result)]
    L.gst_result_free.restype = None
    L.gst_last_error.argtypes = []
    L.gst_last_error.restype = ctypes.c_char_p
    L.gst_version.argtypes = []
    L.gst_version.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(st: int):
    if st != GST_OK:
        raise GstError(st, load().gst_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Result:
    cost: int
    root: int
    edges: np.ndarray            # (n, 4) int32: parent u, child v, weight, parent edge
    status: int = GST_OK
    counters: dict = field(default_factory=dict)
    latency_ms: float = 0.0
    device_ms: float = 0.0

    @property
    def feasible(self) -> bool:
        return self.cost != INFEASIBLE


def _take(r: c_result) -> Result:
    n = r.n_edges
    if n > 0 and r.edges:
        arr = np.ctypeslib.as_array(ctypes.cast(r.edges, ctypes.POINTER(ctypes.c_int32)),
                                    shape=(n * 4,)).reshape(n, 4).copy()
    else:
        arr = np.zeros((0, 4), dtype=np.int32)
    c = r.counters
    out = Result(cost=r.cost, root=r.root, edges=arr, status=r.status,
                 counters={f: getattr(c, f) for f, _ in c_counters._fields_},
                 latency_ms=r.latency_ms, device_ms=r.device_ms)
    load().gst_result_free(ctypes.byref(r))
    return out


class Graph:
    """Device-resident CSR graph plus its group sets (gst_graph)."""

    def __init__(self, rowptr, col, wt):
        L = load()
        self._rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        wt = np.ascontiguousarray(wt, dtype=np.int32)
        self.V = self._rowptr.shape[0] - 1
        self.n_arcs = int(col.shape[0])
        h = ctypes.c_void_p()
        _check(L.gst_graph_create(_ptr(self._rowptr), self.V, _ptr(col), _ptr(wt),
                                  self.n_arcs, ctypes.byref(h)))
        self._h = h
        self.n_groups = 0

    def set_groups(self, gptr, gmem):
        gptr = np.ascontiguousarray(gptr, dtype=np.int64)
        gmem = np.ascontiguousarray(gmem, dtype=np.int32)
        if gmem.shape[0] == 0:
            gmem = np.zeros(1, dtype=np.int32)
        _check(load().gst_groups_set(self._h, _ptr(gptr), _ptr(gmem), gptr.shape[0] - 1))
        self.n_groups = gptr.shape[0] - 1

    def backbone(self):
        """The arcs that queries without a hop bound run on (gst_backbone_get):
        (rowptr int64[V+1], col int32, wt int32)."""
        L = load()
        rp = np.zeros(self.V + 1, dtype=np.int64)
        n = ctypes.c_int64()
        _check(L.gst_backbone_get(self._h, _ptr(rp), None, None, ctypes.byref(n)))
        col = np.zeros(max(n.value, 1), dtype=np.int32)
        wt = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(L.gst_backbone_get(self._h, _ptr(rp), _ptr(col), _ptr(wt), ctypes.byref(n)))
        return rp, col[:n.value], wt[:n.value]

    def set_option(self, key: str, value: int):
        _check(load().gst_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = ctypes.c_int64()
        _check(load().gst_get_option(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def query_bytes(self, k: int, D: int = UNBOUNDED) -> int:
        v = ctypes.c_int64()
        _check(load().gst_query_bytes(self._h, k, D, ctypes.byref(v)))
        return v.value

    def solve(self, group_ids: Sequence[int], D: int = UNBOUNDED) -> Result:
        ids = np.ascontiguousarray(group_ids, dtype=np.int32)
        r = c_result()
        _check(load().gst_solve(self._h, _ptr(ids), ids.shape[0], int(D), ctypes.byref(r)))
        return _take(r)

    def solve_static(self, group_ids: Sequence[int], D: int = UNBOUNDED,
                     edge_cap: int = 4096) -> Result:
        """gst_solve_static: the same answer as solve() with no device
        allocation (state in the library's static scratch); the tree comes
        back through a caller-side buffer of edge_cap records."""
        ids = np.ascontiguousarray(group_ids, dtype=np.int32)
        buf = np.zeros((max(edge_cap, 1), 4), dtype=np.int32)
        r = c_result()
        _check(load().gst_solve_static(self._h, _ptr(ids), ids.shape[0], int(D), _ptr(buf),
                                       int(edge_cap), ctypes.byref(r)))
        c = r.counters
        return Result(cost=r.cost, root=r.root, edges=buf[:r.n_edges].copy(), status=r.status,
                      counters={f: getattr(c, f) for f, _ in c_counters._fields_},
                      latency_ms=r.latency_ms, device_ms=r.device_ms)

    def solve_batch(self, queries: Sequence[Sequence[int]],
                    D: Optional[Sequence[int]] = None) -> List[Result]:
        n = len(queries)
        keep = [np.ascontiguousarray(q, dtype=np.int32) for q in queries]
        qs = (c_query * max(n, 1))()
        for i, a in enumerate(keep):
            qs[i].group_ids = a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            qs[i].k = a.shape[0]
            qs[i].D = UNBOUNDED if D is None else int(D[i])
        rs = (c_result * max(n, 1))()
        st = load().gst_solve_batch(self._h, n, qs, rs)
        out = [_take(rs[i]) for i in range(n)]
        if st != GST_CANCELLED:  # cancelled queries carry their status in the results
            _check(st)
        return out

    def close(self):
        if getattr(self, "_h", None):
            load().gst_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching query server over a resident graph (gst_server)."""

    def __init__(self, graph: Graph, budget_bytes: int = 0, max_slots: int = 0):
        self.graph = graph
        h = ctypes.c_void_p()
        _check(load().gst_server_create(graph._h, int(budget_bytes), int(max_slots),
                                        ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, group_ids: Sequence[int], D: int = UNBOUNDED) -> int:
        a = np.ascontiguousarray(group_ids, dtype=np.int32)
        q = c_query(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), a.shape[0], int(D))
        t = ctypes.c_uint64()
        _check(load().gst_submit(self._h, ctypes.byref(q), ctypes.byref(t)))
        return t.value

    def poll(self, ticket: int) -> Optional[Result]:
        r = c_result()
        st = load().gst_poll(self._h, ticket, ctypes.byref(r))
        if st == GST_NOT_READY:
            return None
        _check(st)
        return _take(r)

    def wait(self, ticket: int) -> Result:
        """The request's result; a cancelled request comes back with status
        GST_CANCELLED (no exception)."""
        r = c_result()
        _check(load().gst_wait(self._h, ticket, ctypes.byref(r)))
        return _take(r)

    def cancel(self, ticket: int):
        _check(load().gst_cancel(self._h, ticket))

    def stat(self, key: str) -> int:
        """Device statistics: "steps", "chunks", "busy_us"."""
        v = ctypes.c_int64()
        _check(load().gst_server_stat(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "_h", None):
            load().gst_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
