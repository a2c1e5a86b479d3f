"""Schedule autotuner for the persistent engine.

Tunables (all runtime options of the C ABI):
* ``chunk``          -- arcs per work item of a query's own frontier rounds
                        (load-balance granularity);
* ``blocks_per_cu``  -- workgroups per CU of the step kernels' grids (0: the
                        occupancy maximum of each kernel);
* ``graph``          -- replay chunks of engine steps from a captured hipGraph;
* ``max_slots``      -- queries and group-distance jobs admitted at once.

``tune()`` times a sample of queries through ``solve_batch`` for every point of
a small grid and returns the fastest; results are kept per (graph, k, D)
bucket in a JSON file (``gpugst_amd/tuned.json`` by default) that ``apply()``
reads back.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from typing import Dict, Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DB = os.path.join(HERE, "tuned.json")

GRID = {
    "chunk": (256, 512, 1024),
    "blocks_per_cu": (0, 2, 3),
    "graph": (0, 1),
    "max_slots": (256, 512),
}
KEYS = ("chunk", "blocks_per_cu", "graph", "max_slots")


def bucket(V: int, arcs: int, k, D: int) -> str:
    """Graphs are bucketed by powers of two of V and of the arc count; k is a
    query size or "mix" for a mixed-k workload."""
    return f"V2^{max(V, 1).bit_length()}_A2^{max(arcs, 1).bit_length()}_k{k}_D{D}"


def load_db(path: str = DEFAULT_DB) -> Dict[str, dict]:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def apply(graph, k, D: int = -1, path: str = DEFAULT_DB) -> Optional[dict]:
    """Set the tuned options for this graph and query shape (k, or "mix"), if
    known."""
    db = load_db(path)
    b = db.get(bucket(graph.V, graph.n_arcs, k, D))
    if not b:
        return None
    for key in KEYS:
        if key in b["best"]:
            graph.set_option(key, int(b["best"][key]))
    return b["best"]


def tune(graph, queries: Sequence, D: int = -1, grid: Optional[dict] = None,
         path: Optional[str] = DEFAULT_DB, repeats: int = 1, log=print, kkey=None) -> dict:
    """Grid search; returns {'best': {...}, 'qps': float, 'trials': [...]}."""
    grid = grid or GRID
    keys = list(grid)
    trials = []
    ref_costs = None
    for combo in itertools.product(*(grid[k] for k in keys)):
        opts = dict(zip(keys, combo))
        for key, val in opts.items():
            graph.set_option(key, val)
        graph.solve_batch(list(queries[: min(len(queries), 8)]), [D] * min(len(queries), 8))
        best_t = None
        for _ in range(repeats):
            t = time.perf_counter()
            res = graph.solve_batch(list(queries), [D] * len(queries))
            dt = time.perf_counter() - t
            best_t = dt if best_t is None else min(best_t, dt)
        costs = [r.cost for r in res]
        if ref_costs is None:
            ref_costs = costs
        elif costs != ref_costs:
            raise AssertionError(f"autotune: costs changed under schedule {opts}")
        qps = len(queries) / best_t
        trials.append(dict(opts, qps=qps))
        log(f"autotune {opts} -> {qps:.1f} q/s")
    best = max(trials, key=lambda t: t["qps"])
    out = {"best": {k: best[k] for k in keys}, "qps": best["qps"], "trials": trials}
    for key in keys:
        graph.set_option(key, out["best"][key])
    if path:
        k = kkey if kkey is not None else len(queries[0])
        db = load_db(path)
        db[bucket(graph.V, graph.n_arcs, k, D)] = out
        with open(path, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return out
