"""Schedule autotuner for the persistent engine.

Tunables (all runtime options of the C ABI):
* ``block``          -- waves per block of the step kernels (2, 4, 8);
* ``blocks_per_cu``  -- workgroups per CU of the step kernels' grids (0: the
                        occupancy maximum of each kernel);
* ``chunk``          -- arcs per work item of a query's own frontier rounds
                        (the frontier-chunk granularity);
* ``merge_tv``       -- live vertices per MERGE / FINAL work item (how many
                        vertices' subset tuples one wave takes per item);
* ``job_lanes``      -- lanes sharing an arc in group-distance job rounds, i.e.
                        how many of a row's dirty chunks (16 labels unbounded,
                        4 hop-bounded) each lane relaxes (0: the smallest power
                        of two giving each lane <= 8);
* ``graph``          -- replay chunks of engine steps from a captured hipGraph;
* ``max_slots``      -- queries and group-distance jobs admitted at once.

``tune()`` times a sample batch through ``solve_batch`` and searches the space
by coordinate descent: starting from the engine defaults it sweeps one knob at
a time over its values, keeps the fastest, and repeats the sweep once more
(about 2 x 14 timed batches instead of the ~2,200 of the full grid).  Every
schedule must return the same costs.  Results are kept per (graph size, k or
"mix", D) bucket in a JSON file (``gpugst_amd/tuned.json`` by default) that
``apply()`` reads back.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DB = os.path.join(HERE, "tuned.json")

GRID = {
    "block": (2, 4, 8),
    "blocks_per_cu": (0, 2, 3),
    "chunk": (256, 512, 1024),
    "merge_tv": (32, 64, 128),
    "job_lanes": (0, 1, 4, 8, 16),
    "graph": (0, 1),
    "max_slots": (256, 512),
}
DEFAULTS = {"block": 2, "blocks_per_cu": 0, "chunk": 512, "merge_tv": 64, "job_lanes": 0, "graph": 0,
            "max_slots": 512}
KEYS = tuple(GRID)


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


def _time(graph, queries, D, opts, repeats):
    for key, val in opts.items():
        graph.set_option(key, val)
    n = len(queries)
    graph.solve_batch(list(queries[: min(n, 8)]), [D] * min(n, 8))  # warm-up (graph capture, arena)
    best_t, res = None, None
    for _ in range(repeats):
        t = time.perf_counter()
        res = graph.solve_batch(list(queries), [D] * n)
        dt = time.perf_counter() - t
        best_t = dt if best_t is None else min(best_t, dt)
    return n / best_t, [r.cost for r in res]


def tune(graph, queries: Sequence, D: int = -1, grid: Optional[dict] = None,
         path: Optional[str] = DEFAULT_DB, repeats: int = 1, log=print, kkey=None,
         sweeps: int = 2) -> dict:
    """Coordinate descent over ``grid``; returns {'best': {...}, 'qps': float,
    'trials': [...]} and leaves the best schedule set on ``graph``."""
    grid = grid or GRID
    keys = list(grid)
    cur = {k: DEFAULTS[k] if DEFAULTS.get(k) in grid[k] else grid[k][0] for k in keys}
    trials = []
    seen = {}
    ref_costs = None

    def run(opts):
        nonlocal ref_costs
        key = tuple(opts[k] for k in keys)
        if key in seen:
            return seen[key]
        qps, costs = _time(graph, queries, D, opts, repeats)
        if ref_costs is None:
            ref_costs = costs
        elif costs != ref_costs:
            raise AssertionError(f"autotune: costs changed under schedule {opts}")
        seen[key] = qps
        trials.append(dict(opts, qps=qps))
        log(f"autotune {opts} -> {qps:.1f} q/s")
        return qps

    best_qps = run(cur)
    for _ in range(sweeps):
        changed = False
        for k in keys:
            for val in grid[k]:
                if val == cur[k]:
                    continue
                opts = dict(cur, **{k: val})
                q = run(opts)
                if q > best_qps:
                    best_qps, cur, changed = q, opts, True
        if not changed:
            break
    out = {"best": dict(cur), "qps": best_qps, "trials": trials, "search": "coordinate descent"}
    for key in keys:
        graph.set_option(key, out["best"][key])
    if path:
        k = kkey if kkey is not None else len(queries[0])
        db = load_db(path)
        db[bucket(graph.V, graph.n_arcs, k, D)] = out
        with open(path, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return out
