"""Data layer: deterministic generators and the reference's file formats."""
import os

import numpy as np
import pytest

from gpugst_amd import data

REF_DATA = "/root/reference/data"


def test_generators_deterministic():
    a = data.make_config(2, scale=0.01, nq=5)
    b = data.make_config(2, scale=0.01, nq=5)
    for x, y in [(a.rowptr, b.rowptr), (a.col, b.col), (a.wt, b.wt), (a.gptr, b.gptr), (a.gmem, b.gmem)]:
        assert np.array_equal(x, y)
    assert all(np.array_equal(p, q) for p, q in zip(a.queries, b.queries))


def test_csr_invariants():
    inst = data.make_config(3, scale=0.01, nq=3)
    rp, col, wt = inst.rowptr, inst.col, inst.wt
    V = inst.V
    assert rp[0] == 0 and rp[-1] == col.shape[0] and np.all(np.diff(rp) >= 0)
    # rows sorted by neighbour, no self loops, weights in range, symmetric
    src = np.repeat(np.arange(V), np.diff(rp))
    assert np.all(src != col)
    assert wt.min() >= 1 and wt.max() <= 100
    fwd = set(zip(src.tolist(), col.tolist(), wt.tolist()))
    assert all((v, u, w) in fwd for u, v, w in list(fwd)[:2000])
    for v in range(0, V, max(1, V // 50)):
        seg = col[rp[v]:rp[v + 1]]
        assert np.all(np.diff(seg) > 0)


def test_group_sizes_and_queries():
    inst = data.make_config(1)
    sizes = np.diff(inst.gptr)
    assert sizes.min() >= 1 and sizes.max() <= 200
    assert 4.0 < sizes.mean() < 6.0  # geometric with mean 5
    for q in inst.queries:
        assert len(set(q.tolist())) == 3
    comp, lcc, size = data.components(inst.rowptr, inst.col)
    assert size == inst.V  # restricted to the largest component


def test_text_and_binary_round_trip(tmp_path):
    inst = data.make_config(1, scale=0.05, nq=4)
    V = inst.V
    src = np.repeat(np.arange(V), np.diff(inst.rowptr))
    m = src < inst.col
    data.write_in(str(tmp_path / "X.in"), V, src[m], inst.col[m], inst.wt[m])
    rp, col, wt = data.read_in(str(tmp_path / "X.in"))
    assert np.array_equal(rp, inst.rowptr) and np.array_equal(col, inst.col) and np.array_equal(wt, inst.wt)
    data.write_bin(str(tmp_path / "X"), inst.rowptr, inst.col, inst.wt)
    rp2, col2, wt2 = data.read_bin(str(tmp_path / "X"))
    assert np.array_equal(rp2, inst.rowptr) and np.array_equal(col2, inst.col) and np.array_equal(wt2, inst.wt)
    data.write_groups(str(tmp_path / "X.g"), inst.gptr, inst.gmem)
    gp, gm = data.read_groups(str(tmp_path / "X.g"))
    assert np.array_equal(gp, inst.gptr) and np.array_equal(gm, inst.gmem)
    data.write_queries(str(tmp_path / "X3.csv"), inst.queries)
    qs = data.read_queries(str(tmp_path / "X3.csv"))
    assert all(np.array_equal(a, b) for a, b in zip(qs, inst.queries))


def test_twitch_excerpt_matches_reference(golden_dir):
    z = np.load(os.path.join(golden_dir, "twitch_excerpt.npz"))
    assert int(z["V"]) == 34118 and int(z["arcs"]) == 858226
    assert z["q3_head"].shape == (16, 3) and z["q5_head"].shape == (16, 5)
    if not os.path.isdir(REF_DATA):
        pytest.skip("reference data not mounted; excerpt checked above")
    bp, col, wt = data.read_bin(os.path.join(REF_DATA, "Twitch"))
    assert col is None and wt is None  # the csr / weight blobs are not shipped
    assert np.array_equal(bp[:256], z["beg_pos_head"])
    qs = data.read_queries(os.path.join(REF_DATA, "Twitch3.csv"))
    assert len(qs) == int(z["n_q3"]) and np.array_equal(np.stack(qs[:16]), z["q3_head"])
    gp, gm = data.read_groups(os.path.join(REF_DATA, "Twitch.g"))
    assert np.array_equal(gm[gp[0]:gp[1]], z["g0"])
    assert np.array_equal(np.diff(gp)[:256], z["sizes_head"])
