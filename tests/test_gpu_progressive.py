"""Progressive answers on the MI355X (options approx_num / approx_den).

The reference's PrunedDP++ returns early once its best feasible tree is within
`maximum_return_app_ratio` of the popped state's cost
(code/PrunedDP++/include/CPUNONHOP.h:499, :730-755), with an int division that
can accept any tree under twice the bound.  Here the test is exact: a query
stops as soon as its cheapest walk-tree costs at most num/den times a certified
lower bound, and reports the bound.  Checked on every answer:
  - the tree is valid (edges exist, edge sum == cost, groups covered, hop bound);
  - OPT <= cost, lower_bound <= OPT and cost * den <= num * lower_bound;
  - ratio 1/1 returns the optimum.
OPT comes from the oracle on the medium graphs and from the exact engine on the
1.1M-vertex graph.
"""
import random

import pytest

import oracle as O
from gpugst_amd import Graph, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu

RATIOS = ((1, 1), (3, 2), (2, 1), (4, 1))


def _check(rp, col, wt, groups, r, opt, num, den, D=-1):
    lb = r.counters["lower_bound"]
    assert r.status == 0
    assert opt <= r.cost and lb <= opt and r.cost * den <= num * lb, (r.cost, lb, opt, num, den)
    if num == den:
        assert r.cost == opt
    assert validate_tree(rp, col, wt, groups, r.cost, r.root, r.edges, D=D)["ok"]


@pytest.mark.parametrize("seed", range(2))
def test_progressive_medium_vs_oracle(seed):
    rng = random.Random(2000 + seed)
    V = 1500
    u, v = data.chung_lu_edges(V, 5 * V, 2.2, 2100 + seed)
    rp, col, wt, _ = data.weighted_csr(V, u, v, 1, 60, 2101 + seed)
    G = 50
    gptr, gmem = data.make_groups(V, G, 4, 30, 2102 + seed)
    g = Graph(rp, col, wt)
    g.set_groups(gptr, gmem)
    qs = [rng.sample(range(G), rng.randint(3, 6)) for _ in range(8)]
    Ds = [-1 if i % 3 else 3 for i in range(len(qs))]
    gl = [[gmem[gptr[x]:gptr[x + 1]] for x in q] for q in qs]
    opt = [O.dpbf(rp, col, wt, gl[i], D=Ds[i])[0] for i in range(len(qs))]
    early = 0
    for num, den in RATIOS:
        g.set_option("approx_num", num)
        g.set_option("approx_den", den)
        print("progressive: seed", seed, "ratio", num, "/", den, flush=True)
        batch = g.solve_batch(qs, Ds)
        for i, q in enumerate(qs):
            if opt[i] == O.INF:
                continue
            for r in (g.solve(q, D=Ds[i]), batch[i]):
                _check(rp, col, wt, gl[i], r, opt[i], num, den, D=Ds[i])
                early += int(r.cost != r.counters["lower_bound"])
    g.close()
    print("progressive: answers below the optimum's certificate", early, flush=True)


def test_progressive_rmat_k5():
    inst = data.make_config(3, nq=1)
    qs = list(data.make_queries(32, 5, inst.meta["eligible"], 3000 + 4 + 10 * 5 + 1))
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    exact = [r.cost for r in g.solve_batch(qs)]
    g.set_option("approx_num", 3)
    g.set_option("approx_den", 2)
    for r, opt, q in zip(g.solve_batch(qs), exact, qs):
        _check(inst.rowptr, inst.col, inst.wt, [inst.group(int(x)) for x in q], r, opt, 3, 2)
    g.close()
