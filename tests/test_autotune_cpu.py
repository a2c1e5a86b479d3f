"""The schedule autotuner's bookkeeping (no GPU): coordinate descent finds
the fastest point of a separable cost, rejects a schedule that changes an
answer, stores results per (graph, k or "mix", D) bucket and applies them
back."""
import os
import time

import pytest

from gpugst_amd import autotune


class FakeResult:
    def __init__(self, cost):
        self.cost = cost


class FakeGraph:
    """solve_batch sleeps according to the options; costs depend on nothing
    unless `bad` names an option value that corrupts them."""
    V, n_arcs = 1000, 6000

    def __init__(self, bad=None):
        self.opts = {}
        self.bad = bad

    def set_option(self, k, v):
        self.opts[k] = v

    def solve_batch(self, qs, Ds=None):
        t = 0.002 * (self.opts.get("chunk", 512) / 256) + 0.001 * self.opts.get("graph", 0)
        t -= 0.0005 * (self.opts.get("blocks_per_cu", 0) == 2)
        t += 0.0008 * (self.opts.get("block", 4) != 8) + 0.0004 * (self.opts.get("merge_tv", 64) != 32)
        time.sleep(max(t, 0.0))
        off = 1 if self.bad and self.opts.get(self.bad[0]) == self.bad[1] else 0
        return [FakeResult(10 + i + off) for i in range(len(qs))]


def test_tune_picks_fastest_and_applies(tmp_path):
    db = os.path.join(tmp_path, "tuned.json")
    g = FakeGraph()
    qs = [[1, 2, 3]] * 200
    grid = {"chunk": (256, 1024), "blocks_per_cu": (0, 2), "graph": (0, 1), "max_slots": (256,),
            "block": (2, 4, 8), "merge_tv": (32, 64, 128)}
    out = autotune.tune(g, qs, grid=grid, path=db, log=lambda *a: None)
    assert out["best"] == {"chunk": 256, "blocks_per_cu": 2, "graph": 0, "max_slots": 256, "block": 8,
                           "merge_tv": 32}
    # coordinate descent: far fewer timed batches than the 72 points of the grid
    assert len(out["trials"]) <= 16
    g2 = FakeGraph()
    assert autotune.apply(g2, 3, -1, path=db) == out["best"]
    assert g2.opts == out["best"]
    assert autotune.apply(g2, 4, -1, path=db) is None  # other k: no entry
    autotune.tune(g, qs, grid=grid, path=db, log=lambda *a: None, kkey="mix")
    assert autotune.apply(FakeGraph(), "mix", -1, path=db) is not None


def test_tune_rejects_schedule_that_changes_costs(tmp_path):
    g = FakeGraph(bad=("graph", 1))
    grid = {"chunk": (256,), "blocks_per_cu": (0,), "graph": (0, 1), "max_slots": (256,)}
    with pytest.raises(AssertionError):
        autotune.tune(g, [[1, 2]] * 20, grid=grid, path=None, log=lambda *a: None)
