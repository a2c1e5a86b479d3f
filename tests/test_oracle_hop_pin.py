"""Independent pins of the oracle's hop-bounded DP (CPU only).

The oracle's hop DP restates the rooted recurrence of
code/D-PrunedDP++/include/CPUHOP.h:192-227 (grow d -> d + 1 while d < D; merge
(v,p1,d1) + (v,p2,d2) with d1 + d2 <= D into height max(d1, d2)).  Two checks
that do not share that recurrence pin it:

  * the center formulation (oracle.hop_center_dp): a tree of diameter <= 2r
    hangs within r hops of a center vertex, one of diameter 2r + 1 within r
    hops of the two ends of a center arc -- equal on every case;
  * brute force over the graph's proper subtrees (oracle.brute_force_hop_trees):
    the DP ranges over tree-like walks, so it is <= the cheapest proper tree of
    diameter <= D, and equal when its own witness is a proper tree.

The last test plants plausible restatement mistakes in a copy of
oracle/gst_oracle.c (compiled with gcc into a temporary directory) and
requires the center check to catch each one.
"""
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from gpugst_amd import data
from gpugst_amd.validate import validate_tree


def _rand_connected(V, E, wmax, rng):
    edges = {}
    for v in range(1, V):
        u = rng.randrange(v)
        edges[(u, v)] = rng.randint(1, wmax)
    while len(edges) < E:
        u, v = rng.randrange(V), rng.randrange(V)
        if u != v:
            edges.setdefault((min(u, v), max(u, v)), rng.randint(1, wmax))
    return [(u, v, w) for (u, v), w in sorted(edges.items())]


def _cases(n, seed, vmax=10, emax=20, kmin=3, kmax=4):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        V = rng.randint(4, vmax)
        el = _rand_connected(V, rng.randint(V - 1, min(emax, V * (V - 1) // 2)), 9, rng)
        k = rng.randint(kmin, kmax)
        groups = [sorted(rng.sample(range(V), rng.randint(1, 3))) for _ in range(k)]
        out.append((V, el, groups))
    return out


def _center_mismatches(cases, Ds, library=None):
    bad = []
    for V, el, groups in cases:
        rp, col, wt = data.csr_from_edges(V, el)
        for D in Ds:
            c = O.dpbf(rp, col, wt, groups, D=D, library=library)[0]
            cc = O.hop_center_dp(V, el, groups, D)
            if c != cc:
                bad.append((V, el, groups, D, c, cc))
    return bad


def test_center_formulation_small_cases():
    """Hand-checked: a path a-b-c-d (weights 1) with groups {a}, {d}: diameter
    3 is needed, so D <= 2 is infeasible and D = 3 costs 3; a star whose
    center is in no group still reaches three leaves within D = 2."""
    path = [(0, 1, 1), (1, 2, 1), (2, 3, 1)]
    assert [O.hop_center_dp(4, path, [[0], [3]], D) for D in range(5)] == [O.INF, O.INF, O.INF, 3, 3]
    star = [(0, 1, 2), (0, 2, 3), (0, 3, 4)]
    assert O.hop_center_dp(4, star, [[1], [2], [3]], 1) == O.INF
    assert O.hop_center_dp(4, star, [[1], [2], [3]], 2) == 9
    assert O.brute_force_hop_trees(4, star, [[1], [2], [3]], 3) == [O.INF, O.INF, 9, 9]


def test_hop_dp_equals_center_formulation():
    cases = _cases(80, 31) + _cases(30, 32, kmin=2, kmax=5)
    assert _center_mismatches(cases, range(0, 7)) == []


def test_hop_dp_against_proper_trees():
    cases = _cases(60, 33, vmax=9, emax=13)
    equal = 0
    proper = 0
    for V, el, groups in cases:
        rp, col, wt = data.csr_from_edges(V, el)
        bf = O.brute_force_hop_trees(V, el, groups, 5)
        for D in range(0, 6):
            c, root, edges, _ = O.dpbf(rp, col, wt, groups, D=D)
            assert c <= bf[D], (V, el, groups, D, c, bf[D])
            if c == O.INF:
                continue
            v = validate_tree(rp, col, wt, groups, c, root, np.array(edges).reshape(-1, 4), D=D)
            assert v["ok"], v
            if not v["walk"]:
                proper += 1
                assert c == bf[D], (V, el, groups, D, c, bf[D])
            equal += int(c == bf[D])
    # the equality branch is the one that catches a too-strict DP: it must run
    assert proper >= 200 and equal >= proper


def test_golden_hop_values_pinned(golden_dir):
    """tiny.json's hop values (written by the oracle) agree with the center
    formulation, so they pin the GPU engine independently of the oracle."""
    import json
    cases = json.load(open(os.path.join(golden_dir, "tiny.json")))
    for c in cases:
        for D, hc in c["hop"].items():
            assert O.hop_center_dp(c["V"], c["edges"], c["groups"], int(D)) == hc


# (pattern in gst_oracle.c, planted mistake)
MUTANTS = [
    ("int32_t nd = d > d2 ? d : d2;", "int32_t nd = d + d2;"),          # merge height: sum, not max
    ("if (hop && d + d2 > D) break;", "if (hop && d2 > D) break;"),    # merge bound ignores d1
    ("if (!hop || d < D) {", "if (!hop || d + 1 < D) {"),              # grow one hop short
    ("if (hop && d + d2 > D) break;", "if (hop && d + d2 >= D) break;"),  # d1 + d2 = D excluded
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("mi", range(len(MUTANTS)))
def test_planted_mistakes_are_caught(tmp_path, mi):
    src = open(os.path.join(os.path.dirname(O.__file__), "gst_oracle.c")).read()
    old, new = MUTANTS[mi]
    assert src.count(old) == 1, old
    mpath = tmp_path / "mutant.c"
    mpath.write_text(src.replace(old, new))
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-o", str(so), str(mpath)], check=True)
    lib = O.load_variant(str(so))
    bad = _center_mismatches(_cases(80, 31), range(0, 7), library=lib)
    assert bad, f"planted mistake {new!r} not caught"


def test_reference_instance_hop_values_pinned(golden_dir):
    """The reference's own instance (code/GPUGST/simple_iterative_tests.text,
    V = 10,000): the committed hop optima (infeasible up to D = 5, 18 at
    D = 6) equal the center formulation."""
    z = np.load(os.path.join(golden_dir, "simple_iterative_tests.npz"))
    rp, col, wt = z["rowptr"], z["col"], z["wt"]
    V = rp.shape[0] - 1
    src = np.repeat(np.arange(V), np.diff(rp))
    m = src < col
    el = list(zip(src[m].tolist(), col[m].tolist(), wt[m].tolist()))
    gl = [z["gmem"][z["gptr"][i]:z["gptr"][i + 1]].tolist() for i in range(len(z["gptr"]) - 1)]
    for D, cost in zip(z["D"], z["cost"]):
        if int(D) >= 0:
            assert O.hop_center_dp(V, el, gl, int(D)) == int(cost), int(D)
