"""CPU build of the engine's device code (test infrastructure).

``gpugst_amd/csrc/gst_device.hip`` is compiled with g++ against the SIMT
emulation layer in ``hip/hip_runtime.h`` (64 host threads per wave), so the
same item and plan functions run on the host.  Used by CPU tests to check the
device logic (termination, bounds, exact costs) where no GPU exists.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "_build", "libgst_sim.so")
_lib = None


def build(debug_flags=("-O1", "-g")):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    src = os.path.join(HERE, "sim_engine.cpp")
    cmd = ["g++", "-std=c++20", *debug_flags, "-fPIC", "-shared", "-pthread", "-DGST_CPU_SIM",
           "-x", "c++", "-I", HERE, "-I", os.path.join(ROOT, "gpugst_amd", "csrc"), src, "-o", LIB]
    subprocess.run(cmd, check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        dev = os.path.join(ROOT, "gpugst_amd", "csrc", "gst_device.hip")
        if (not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(dev)
                or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "sim_engine.cpp"))):
            build()
        L = ctypes.CDLL(LIB)
        L.gst_sim_solve.restype = ctypes.c_int
        _lib = L
    return _lib


def reverse_arcs(rowptr, col, wt):
    V = rowptr.shape[0] - 1
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(rowptr))
    key_fwd = (src * (V + 1) + col) * (1 << 21) + wt
    key_rev = (col.astype(np.int64) * (V + 1) + src) * (1 << 21) + wt
    order = np.argsort(key_fwd, kind="stable")
    pos = np.searchsorted(key_fwd[order], key_rev)
    return order[pos].astype(np.int32)


def solve(rowptr, col, wt, gptr, gmem, qgroups, D=-1, edge_cap=4096, chunk=512, n_waves=2,
          max_steps=100000, verbose=0):
    P = ctypes.c_void_p
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    wt = np.ascontiguousarray(wt, np.int32)
    rev = reverse_arcs(rowptr, col, wt)
    gptr = np.ascontiguousarray(gptr, np.int64)
    gmem = np.ascontiguousarray(gmem if len(gmem) else np.zeros(1), np.int32)
    q = np.ascontiguousarray(qgroups, np.int32)
    out = np.zeros(4, np.int32)
    edges = np.zeros(edge_cap * 4, np.int32)
    ctr = np.zeros(6, np.int64)
    p = lambda a: a.ctypes.data_as(P)
    rc = lib().gst_sim_solve(p(rowptr), ctypes.c_int64(rowptr.shape[0] - 1), p(col), p(wt),
                             ctypes.c_int64(col.shape[0]), p(rev), p(gptr), p(gmem), p(q),
                             ctypes.c_int32(q.shape[0]), ctypes.c_int32(D), ctypes.c_int32(edge_cap),
                             ctypes.c_int32(chunk), ctypes.c_int32(n_waves), ctypes.c_int32(max_steps),
                             ctypes.c_int32(verbose), p(out), p(edges), p(ctr))
    n = int(out[2])
    return dict(rc=rc, cost=int(out[0]), root=int(out[1]), edges=edges[:4 * n].reshape(n, 4),
                steps=int(out[3]), counters=ctr)
