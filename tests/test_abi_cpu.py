"""CPU-side checks of the native boundary: the library loads, exports every
symbol include/gst.h declares, and rejects malformed input before touching the
device (no compute calls are made here)."""
import ctypes
import re
import os

import numpy as np
import pytest

from gpugst_amd import capi
from gpugst_amd import build as b


def test_header_declares_exactly_the_exports():
    hdr = open(os.path.join(b.ROOT, "include", "gst.h")).read()
    declared = set(re.findall(r"\b(gst_[a-z_]+)\s*\(", hdr))
    assert declared == set(capi.EXPORTED), declared ^ set(capi.EXPORTED)


def test_library_loads_and_exports():
    L = capi.load()
    for name in capi.EXPORTED:
        assert hasattr(L, name), name
    assert b"gfx950" in L.gst_version()


def _create(rowptr, col, wt):
    L = capi.load()
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    wt = np.ascontiguousarray(wt, np.int32)
    h = ctypes.c_void_p()
    st = L.gst_graph_create(rowptr.ctypes.data_as(ctypes.c_void_p), rowptr.shape[0] - 1,
                            col.ctypes.data_as(ctypes.c_void_p), wt.ctypes.data_as(ctypes.c_void_p),
                            col.shape[0], ctypes.byref(h))
    return st, L.gst_last_error().decode()


@pytest.mark.parametrize("case", ["bad_start", "decreasing", "col_range", "weight_zero",
                                  "weight_big", "asymmetric", "asym_weight"])
def test_graph_create_rejects_malformed_input(case):
    rp = [0, 1, 2]
    col = [1, 0]
    wt = [5, 5]
    if case == "bad_start":
        rp = [1, 1, 2]
    elif case == "decreasing":
        rp = [0, 2, 1]
    elif case == "col_range":
        col = [2, 0]
    elif case == "weight_zero":
        wt = [0, 0]
    elif case == "weight_big":
        wt = [1 << 21, 1 << 21]
    elif case == "asymmetric":
        rp, col, wt = [0, 1, 1], [1], [5]
    elif case == "asym_weight":
        wt = [5, 6]
    st, msg = _create(rp, col, wt)
    assert st == capi.GST_INVALID_ARG, (case, st, msg)
    assert msg


def test_null_and_range_arguments():
    L = capi.load()
    r = capi.c_result()
    assert L.gst_solve(None, None, 3, -1, ctypes.byref(r)) == capi.GST_INVALID_ARG
    assert L.gst_set_option(None, b"chunk", 5) == capi.GST_INVALID_ARG
    assert L.gst_server_create(None, 0, 0, ctypes.byref(ctypes.c_void_p())) == capi.GST_INVALID_ARG
    L.gst_result_free(None)
    L.gst_result_free(ctypes.byref(r))
