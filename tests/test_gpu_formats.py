"""GPU parity from the reference's on-disk formats.

The 20,000-vertex configuration is written as the reference stores a dataset
-- int64 X_beg_pos.bin / X_csr.bin / X_weight.bin (code/GPUGST+/include/
graph.hpp:12-118), the edge list X.in (code/GPUGST/include/
graph_hash_of_mixed_weighted_read_for_GSTP.h:75-148), the groups X.g (:190-239)
and the queries X3.csv (:240-286) -- then read back by the native readers and
solved: every cost equals the committed fixture, every tree is valid.
"""
import json
import os

import numpy as np
import pytest

from gpugst_amd import Graph, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def test_c1_from_reference_files(tmp_path, golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "c1_costs.json")))
    inst = data.make_config(1)
    prefix = str(tmp_path / "C1")
    data.write_bin(prefix, inst.rowptr, inst.col, inst.wt)
    V = inst.V
    src = np.repeat(np.arange(V), np.diff(inst.rowptr))
    m = src < inst.col
    data.write_in(prefix + ".in", V, src[m], inst.col[m], inst.wt[m])
    data.write_groups(prefix + ".g", inst.gptr, inst.gmem)
    data.write_queries(prefix + "3.csv", inst.queries)

    rp, col, wt = data.read_bin(prefix)
    gptr, gmem = data.read_groups(prefix + ".g")
    qs = data.read_queries(prefix + "3.csv")
    assert len(qs) == len(ref["costs"])
    g = Graph(rp, col, wt)
    g.set_groups(gptr, gmem)
    print("formats: batch of", len(qs), "from the binary CSR", flush=True)
    rs = g.solve_batch(qs)
    for i, (q, r) in enumerate(zip(qs, rs)):
        assert r.cost == ref["costs"][i], (i, r.cost, ref["costs"][i])
        groups = [gmem[gptr[x]:gptr[x + 1]] for x in q]
        assert validate_tree(rp, col, wt, groups, r.cost, r.root, r.edges)["ok"], i
    g.close()

    rp2, col2, wt2 = data.read_in(prefix + ".in")
    g = Graph(rp2, col2, wt2)
    g.set_groups(gptr, gmem)
    for i in range(0, len(qs), 10):
        print("formats: query", i, "from the edge list", flush=True)
        assert g.solve(qs[i]).cost == ref["costs"][i]
    g.close()
