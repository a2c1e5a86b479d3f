"""Writes the committed golden fixtures under tests/golden/.

Run from the repository root:  python tests/golden/make_golden.py

* simple_iterative_tests.npz -- the reference's only complete instance
  (code/GPUGST/simple_iterative_tests.text: V=10,000, 50,000 edges, 4 query
  groups; it ships with no recorded answer).  Stored as CSR + groups + the
  optimum for D in {-1, 1, ..., 6}, each computed by the oracle and the
  unbounded one cross-checked by Dreyfus-Wagner on the GST -> Steiner
  reduction.  Needs /root/reference (read here only, never at test time).
* tiny.json -- random graphs with V <= 11 whose unbounded optimum is pinned
  by brute force over vertex subsets AND the Dreyfus-Wagner reduction AND the
  oracle, plus hop-bounded optima (D = 0..4) from the oracle, each equal to the
  center formulation (oracle.hop_center_dp) and no more than the cheapest
  proper subtree of that diameter (oracle.brute_force_hop_trees).
* c1_costs.json -- optimal costs of the 50 queries of the V=20k Chung-Lu
  configuration, from the oracle, with a checksum of the generated instance so
  generator drift is caught.
* twitch_excerpt.npz -- a slice of the reference's Twitch data files (format
  fixture for the readers).
"""
from __future__ import annotations

import hashlib
import json
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from gpugst_amd import data  # noqa: E402

REF = "/root/reference"


def instance_digest(inst) -> str:
    h = hashlib.sha256()
    for a in (inst.rowptr, inst.col, inst.wt, inst.gptr, inst.gmem):
        h.update(np.ascontiguousarray(a).tobytes())
    for q in inst.queries:
        h.update(np.ascontiguousarray(q, np.int32).tobytes())
    return h.hexdigest()


def rand_connected(V, E, wmax, rng):
    edges = {}
    for v in range(1, V):
        u = rng.randrange(v)
        edges[(u, v)] = rng.randint(1, wmax)
    while len(edges) < E:
        u, v = rng.randrange(V), rng.randrange(V)
        if u != v:
            edges.setdefault((min(u, v), max(u, v)), rng.randint(1, wmax))
    return [(u, v, w) for (u, v), w in sorted(edges.items())]


def make_tiny(n=60, seed=7):
    rng = random.Random(seed)
    cases = []
    for _ in range(n):
        V = rng.randint(2, 11)
        E = rng.randint(V - 1, min(V * (V - 1) // 2, 18))
        el = rand_connected(V, E, 9, rng)
        k = rng.randint(1, min(5, V))
        groups = [sorted(rng.sample(range(V), rng.randint(1, min(3, V)))) for _ in range(k)]
        rp, col, wt = data.csr_from_edges(V, el)
        c_or, *_ = O.dpbf(rp, col, wt, groups)
        c_bf = O.brute_force_gst(V, el, groups)
        c_dw = O.gst_via_steiner(rp, col, wt, groups)
        assert c_or == c_bf == c_dw, (el, groups, c_or, c_bf, c_dw)
        hop = {str(D): O.dpbf(rp, col, wt, groups, D=D)[0] for D in range(0, 5)}
        # the hop values are pinned by the center formulation, which does not
        # share the oracle's rooted recurrence, and bounded by proper trees
        bf = O.brute_force_hop_trees(V, el, groups, 4)
        for D, hc in hop.items():
            assert hc == O.hop_center_dp(V, el, groups, int(D)), (el, groups, D)
            assert hc <= bf[int(D)], (el, groups, D)
        cases.append(dict(V=V, edges=el, groups=groups, cost=c_or, hop=hop))
    return cases


def main():
    # tiny brute-force-pinned cases
    tiny = make_tiny()
    with open(os.path.join(HERE, "tiny.json"), "w") as f:
        json.dump(tiny, f)
    print("tiny.json", len(tiny))

    # C1 optimal costs
    inst = data.make_config(1)
    costs = []
    for qi in range(len(inst.queries)):
        c, *_ = O.dpbf(inst.rowptr, inst.col, inst.wt, inst.query_groups(qi))
        costs.append(int(c))
    with open(os.path.join(HERE, "c1_costs.json"), "w") as f:
        json.dump(dict(digest=instance_digest(inst), queries=[q.tolist() for q in inst.queries],
                       costs=costs), f)
    print("c1_costs.json", costs[:10])

    if not os.path.isdir(REF):
        print("reference not mounted: keeping the existing reference-derived fixtures")
        return
    # the reference's GSTP instance
    V, edges, groups, order = data.read_gstp(
        os.path.join(REF, "code/GPUGST/simple_iterative_tests.text"))
    rp, col, wt = data.csr_from_edges(V, edges)
    gl = [groups[g] for g in sorted(groups)]
    opt = {}
    for D in [-1, 1, 2, 3, 4, 5, 6]:
        opt[D] = O.dpbf(rp, col, wt, gl, D=D)[0]
    dw = O.gst_via_steiner(rp, col, wt, gl)
    assert dw == opt[-1], (dw, opt)
    gptr = np.zeros(len(gl) + 1, np.int64)
    for i, g in enumerate(gl):
        gptr[i + 1] = gptr[i] + len(g)
    np.savez_compressed(os.path.join(HERE, "simple_iterative_tests.npz"), rowptr=rp, col=col, wt=wt,
                        gptr=gptr, gmem=np.concatenate([np.asarray(g, np.int32) for g in gl]),
                        D=np.asarray(sorted(opt), np.int32),
                        cost=np.asarray([opt[d] for d in sorted(opt)], np.int32))
    print("simple_iterative_tests", opt)

    # Twitch excerpt (format fixture)
    bp = np.fromfile(os.path.join(REF, "data/Twitch_beg_pos.bin"), dtype=np.int64)
    q3 = data.read_queries(os.path.join(REF, "data/Twitch3.csv"))
    q5 = data.read_queries(os.path.join(REF, "data/Twitch5.csv"))
    gptr_t, gmem_t = data.read_groups(os.path.join(REF, "data/Twitch.g"))
    np.savez_compressed(os.path.join(HERE, "twitch_excerpt.npz"),
                        beg_pos_head=bp[:256], V=np.int64(bp.shape[0] - 1), arcs=bp[-1],
                        q3_head=np.stack(q3[:16]), q5_head=np.stack(q5[:16]),
                        n_groups=np.int64(gptr_t.shape[0] - 1), g0=gmem_t[gptr_t[0]:gptr_t[1]],
                        sizes_head=np.diff(gptr_t)[:256], n_q3=np.int64(len(q3)))
    print("twitch excerpt", bp.shape, len(q3), gptr_t.shape)


if __name__ == "__main__":
    main()
