"""hipGraph replay of the engine steps (option graph=1) against eager launches.

The captured graph holds kernel launches only (lean, heavy, plan per step, 16
steps); every epoch replays it until the plan kernel ends the epoch.  A batch
runs through several epochs and many replays; it is run three times with
graph=1, with eager batches and synchronisations in between, and every replay
result must equal the eager one (cost, root and tree validity).
"""
import pytest

import oracle as O
from gpugst_amd import Graph, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager():
    inst = data.make_config(2, scale=0.02, nq=8)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    qs = list(inst.queries)
    Ds = [-1 if i % 4 else 3 for i in range(len(qs))]
    g.set_option("graph", 0)
    eager = g.solve_batch(qs, Ds)
    for i, r in enumerate(eager):
        assert r.cost == O.dpbf(inst.rowptr, inst.col, inst.wt, inst.query_groups(i), D=Ds[i])[0]
    for rep in range(3):
        print("graph replay batch", rep, flush=True)
        g.set_option("graph", 1)
        rg = g.solve_batch(qs, Ds)
        g.set_option("graph", 0)
        re = g.solve_batch(qs[:3], Ds[:3])  # eager work between the replays
        for i in range(3):
            assert re[i].cost == eager[i].cost
        for i, (a, b) in enumerate(zip(rg, eager)):
            assert a.status == 0 and a.cost == b.cost, (rep, i, a.cost, b.cost)
            if a.feasible:
                v = validate_tree(inst.rowptr, inst.col, inst.wt, inst.query_groups(i), a.cost, a.root,
                                  a.edges, D=Ds[i])
                assert v["ok"], (rep, i, v)
    g.close()


@pytest.mark.parametrize("opts", [{"block": 2}, {"block": 8}, {"merge_tv": 32}, {"merge_tv": 128},
                                  {"job_lanes": 4}, {"job_lanes": 16}, {"block": 8, "graph": 1}])
def test_schedules_equal_default(opts):
    """The autotuner's knobs (waves per block, merge work-item size, lanes per
    arc in job rounds, graph replay) change how work is cut, never the
    answers: every schedule returns the default schedule's costs, with valid
    trees."""
    inst = data.make_config(2, scale=0.02, nq=8)
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    qs = list(inst.queries)
    Ds = [-1 if i % 4 else 3 for i in range(len(qs))]
    base = g.solve_batch(qs, Ds)
    for k, v in opts.items():
        g.set_option(k, v)
    print("schedule", opts, flush=True)
    got = g.solve_batch(qs, Ds)
    single = g.solve(qs[1], D=Ds[1])
    assert single.cost == base[1].cost
    for i, (a, b) in enumerate(zip(got, base)):
        assert a.status == 0 and a.cost == b.cost, (opts, i, a.cost, b.cost)
        if a.feasible:
            assert validate_tree(inst.rowptr, inst.col, inst.wt, inst.query_groups(i), a.cost, a.root,
                                 a.edges, D=Ds[i])["ok"]
    g.close()
