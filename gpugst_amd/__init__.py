"""gpugst_amd -- MI355X-native optimal group Steiner tree engine.

The product is the native library ``_lib/libgst.so`` (HIP kernels for gfx950
behind the C ABI of ``include/gst.h``).  This package adds the Python host
mirror of that ABI (:mod:`gpugst_amd.capi`), the data layer for the reference's
formats and the synthetic benchmark inputs (:mod:`gpugst_amd.data`) and a tree
validator (:mod:`gpugst_amd.validate`).
"""
from .capi import (INFEASIBLE, UNBOUNDED, MAX_D, MAX_K, Graph, GstError, Result, Server,
                   load)
from .validate import validate_tree

__all__ = ["INFEASIBLE", "UNBOUNDED", "MAX_D", "MAX_K", "Graph", "GstError", "Result",
           "Server", "load", "validate_tree"]
