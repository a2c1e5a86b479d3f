"""gst_solve_static on the MI355X: the allocation-free variant of gst_solve
(state in the library's static device scratch) returns gst_solve's answer.

Two calls in a row, the second with a smaller input: a query of the
20,000-vertex configuration, then the tiny fixtures (smaller graphs, every D)
through the same scratch.  Costs are bit-exact against gst_solve and the
committed fixtures; every tree is validated.  A query whose state exceeds the
scratch is rejected before anything runs, and gst_solve still answers it.
"""
import json
import os

import numpy as np
import pytest

from gpugst_amd import Graph, GstError, data
from gpugst_amd.capi import GST_OUT_OF_MEMORY
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _groups_csr(groups):
    gptr = np.zeros(len(groups) + 1, np.int64)
    for i, m in enumerate(groups):
        gptr[i + 1] = gptr[i] + len(m)
    return gptr, np.concatenate([np.asarray(m, np.int32) for m in groups])


def _same(rp, col, wt, groups, r, p, expect, D=-1):
    assert r.status == 0 and p.status == 0
    assert r.cost == p.cost == expect, (r.cost, p.cost, expect)
    if r.feasible:
        assert validate_tree(rp, col, wt, groups, r.cost, r.root, r.edges, D=D)["ok"]


def test_static_equals_solve(golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "c1_costs.json")))
    inst = data.make_config(1)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    for qi in (0, 1):
        _say("static: 20k-vertex configuration, query", qi)
        r = g.solve_static(inst.queries[qi])
        p = g.solve(inst.queries[qi])
        _same(inst.rowptr, inst.col, inst.wt, inst.query_groups(qi), r, p, ref["costs"][qi])
    g.close()
    # the second input is smaller: tiny graphs, every bound
    cases = json.load(open(os.path.join(golden_dir, "tiny.json")))
    for ci in range(0, len(cases), 6):
        c = cases[ci]
        _say("static: tiny case", ci, "V", c["V"], "k", len(c["groups"]))
        rp, col, wt = data.csr_from_edges(c["V"], c["edges"])
        t = Graph(rp, col, wt)
        t.set_groups(*_groups_csr(c["groups"]))
        ids = list(range(len(c["groups"])))
        for D, expect in [(-1, c["cost"])] + [(int(d), hc) for d, hc in c["hop"].items()]:
            _same(rp, col, wt, c["groups"], t.solve_static(ids, D=D), t.solve(ids, D=D), expect, D=D)
        t.close()


def test_static_rejects_oversized_state():
    # a path of 200,000 vertices: a k = 6 query needs ~100 MiB of states
    V = 200_000
    u = np.arange(V - 1)
    rp, col, wt = data.csr_from_edges(V, list(zip(u.tolist(), (u + 1).tolist(), [1] * (V - 1))))
    g = Graph(rp, col, wt)
    groups = [[i * 1000] for i in range(6)]
    g.set_groups(*_groups_csr(groups))
    _say("static: oversized query is rejected")
    with pytest.raises(GstError) as ei:
        g.solve_static(list(range(6)))
    assert ei.value.status == GST_OUT_OF_MEMORY
    _say("static: gst_solve answers it")
    assert g.solve(list(range(6))).cost == 5000
    g.close()
