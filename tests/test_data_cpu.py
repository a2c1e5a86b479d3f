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


def test_native_readers_edge_cases(tmp_path):
    # .in: first field of the first line is V; only 3-field lines are edges;
    # blanks of any kind separate fields; CRLF endings
    p = tmp_path / "e.in"
    p.write_bytes(b"5 3\n0 1 7\r\n1  2\t9 \n\n3 4\n2 4 1\n")
    V, u, v, w = data.read_edges(str(p))
    assert V == 5 and u.tolist() == [0, 1, 2] and v.tolist() == [1, 2, 4] and w.tolist() == [7, 9, 1]
    # a vertex id outside [0, V) or a non-integer field is an error with its line
    p.write_bytes(b"3 1\n0 1 2\n0 5 1\n")
    with pytest.raises(ValueError, match="line 3"):
        data.read_edges(str(p))
    p.write_bytes(b"3 1\n0 1 x\n")
    with pytest.raises(ValueError, match="line 2"):
        data.read_edges(str(p))
    # .g: "gN:" is group N-1 whatever the line order; members sorted; a
    # trailing blank and missing groups are fine
    g = tmp_path / "x.g"
    g.write_text("g3:9 4 \ng1:2 1\n\ng5:7\n")
    gp, gm = data.read_groups(str(g))
    assert gp.tolist() == [0, 2, 2, 4, 4, 5] and gm.tolist() == [1, 2, 4, 9, 7]
    g.write_text("g1:1 2\nh2:3\n")
    with pytest.raises(ValueError, match="line 2"):
        data.read_groups(str(g))
    q = tmp_path / "x3.csv"
    q.write_text("1 2 3 \n\n4 5 6\n")
    assert [a.tolist() for a in data.read_queries(str(q))] == [[1, 2, 3], [4, 5, 6]]
    with pytest.raises(OSError):
        data.read_queries(str(tmp_path / "missing.csv"))


def test_native_reader_large_file(tmp_path):
    # 3M edges in a few seconds, multithreaded, identical to the arrays written
    rng = np.random.default_rng(3)
    V, m = 1_000_000, 3_000_000
    u = rng.integers(0, V, m).astype(np.int32)
    v = rng.integers(0, V, m).astype(np.int32)
    w = rng.integers(1, 1000, m).astype(np.int32)
    p = str(tmp_path / "big.in")
    with open(p, "w") as f:
        f.write(f"{V} {m}\n")
        np.savetxt(f, np.stack([u, v, w], 1), fmt="%d")
    import time
    t = time.time()
    V2, u2, v2, w2 = data.read_edges(p)
    assert time.time() - t < 10
    assert V2 == V and np.array_equal(u2, u) and np.array_equal(v2, v) and np.array_equal(w2, w)
