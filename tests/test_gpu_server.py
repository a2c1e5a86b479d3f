"""Serving: continuous batching inside one persistent engine.

Requests that arrive while others run are admitted as slots and memory free
up; each result must be the same alone and in the mix (cost bit-exact, tree
valid), whatever else shares the device.
"""
import threading
import time

import numpy as np
import pytest

import oracle as O
from gpugst_amd import Graph, Server, data
from gpugst_amd.validate import validate_tree

pytestmark = pytest.mark.gpu


def _instance():
    inst = data.make_config(2, scale=0.02, nq=1)
    rng = np.random.default_rng(5)
    elig = inst.meta["eligible"]
    qs, Ds = [], []
    for i in range(24):
        k = int(rng.integers(2, 7))
        qs.append(data.make_queries(1, k, elig, 777 + i)[0])
        Ds.append(-1 if i % 4 else 3)
    return inst, qs, Ds


def test_mix_equals_alone():
    inst, qs, Ds = _instance()
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    alone = [g.solve(q, D=d) for q, d in zip(qs, Ds)]
    # a small arena forces admission to wait for memory
    per = max(g.query_bytes(len(q), d) for q, d in zip(qs, Ds))
    srv = Server(g, budget_bytes=4 * per, max_slots=8)
    tickets = [None] * len(qs)

    def submitter():
        for i, (q, d) in enumerate(zip(qs, Ds)):
            tickets[i] = srv.submit(q, D=d)
            time.sleep(0.002 * (i % 3))

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    mixed = [srv.wait(t) for t in tickets]
    srv.close()
    for i, (a, b) in enumerate(zip(alone, mixed)):
        print("query", i, "k", len(qs[i]), "D", Ds[i], "alone", a.cost, "mix", b.cost, flush=True)
        assert a.cost == b.cost
        if i % 6 == 0:
            groups = inst.query_groups(0)
            groups = [inst.group(int(x)) for x in qs[i]]
            assert O.dpbf(inst.rowptr, inst.col, inst.wt, groups, D=Ds[i])[0] == a.cost
        if b.feasible:
            groups = [inst.group(int(x)) for x in qs[i]]
            v = validate_tree(inst.rowptr, inst.col, inst.wt, groups, b.cost, b.root, b.edges, D=Ds[i])
            assert v["ok"], v
    g.close()


def test_admission_settings_and_stats():
    """The server steps continuously; whether admission takes every request at
    once or holds requests back to share a job, each answer equals the
    one-query call, and the server reports its engine steps and device time."""
    inst, qs, Ds = _instance()
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    alone = [g.solve(q, D=d).cost for q, d in zip(qs, Ds)]
    for wait_us, batch, chunk in ((0, 1, 1), (20000, 8, 4)):
        g.set_option("serve_wait_us", wait_us)
        g.set_option("serve_batch", batch)
        g.set_option("serve_chunk", chunk)
        srv = Server(g)
        tickets = [srv.submit(q, D=d) for q, d in zip(qs, Ds)]
        mixed = [srv.wait(t).cost for t in tickets]
        steps, chunks, busy = srv.stat("steps"), srv.stat("chunks"), srv.stat("busy_us")
        srv.close()
        print("wait_us", wait_us, "batch", batch, "steps", steps, "chunks", chunks, "busy_us", busy, flush=True)
        assert mixed == alone
        assert steps >= chunks > 0 and busy > 0
    g.close()


def test_poll_and_unknown_ticket():
    inst, qs, Ds = _instance()
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    srv = Server(g, max_slots=4)
    t = srv.submit(qs[0], D=-1)
    r = None
    t0 = time.time()
    while r is None and time.time() - t0 < 60:
        r = srv.poll(t)
        time.sleep(0.001)
    assert r is not None and r.cost == g.solve(qs[0]).cost
    with pytest.raises(Exception):
        srv.poll(123456789)
    srv.close()
    g.close()


def test_deadline_and_cancel_on_device():
    """Cancellation polled by the plan kernel: with a 1 ms deadline the
    queries of a batch stop on the device (status CANCELLED) or finish with
    the exact cost; the engine then answers exactly again.  A cancelled server
    request comes back CANCELLED or exact; destroying a server with requests
    in flight wakes a waiting thread."""
    from gpugst_amd.capi import GST_CANCELLED
    inst, qs, Ds = _instance()
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    exact = [g.solve(q, D=d).cost for q, d in zip(qs, Ds)]
    g.set_option("deadline_ms", 1)
    rs = g.solve_batch(qs * 4, Ds * 4)
    g.set_option("deadline_ms", 0)
    n_c = sum(r.status == GST_CANCELLED for r in rs)
    print("deadline: cancelled", n_c, "of", len(rs), flush=True)
    assert n_c > 0
    for i, r in enumerate(rs):
        assert r.status == GST_CANCELLED or r.cost == exact[i % len(qs)], i
    assert [r.cost for r in g.solve_batch(qs, Ds)] == exact
    srv = Server(g)
    t = srv.submit(qs[0], D=Ds[0])
    srv.cancel(t)
    r = srv.wait(t)
    assert r.status == GST_CANCELLED or r.cost == exact[0]
    tickets = [srv.submit(q, D=d) for q, d in zip(qs, Ds)]
    got = []
    th = threading.Thread(target=lambda: got.append(srv.wait(tickets[-1])))
    th.start()
    srv.close()
    th.join(timeout=60)
    assert not th.is_alive() and len(got) == 1
    assert got[0].status == GST_CANCELLED or got[0].cost == exact[-1]
    g.close()


def test_finite_state_count():
    inst, qs, Ds = _instance()
    g = Graph(inst.rowptr, inst.col, inst.wt)
    g.set_groups(inst.gptr, inst.gmem)
    q = qs[1]
    r0 = g.solve(q)
    g.set_option("count_states", 1)
    r1 = g.solve(q)
    rb = g.solve_batch([q, q])
    assert r0.counters["states_finite"] == -1
    assert r1.cost == r0.cost == rb[0].cost
    # the singleton states of every vertex the groups reach are finite
    reach = int(np.count_nonzero(O.sssp(inst.rowptr, inst.col, inst.wt, inst.group(int(q[0]))) < O.INF))
    for r in (r1, rb[0]):
        assert r.counters["states_finite"] >= reach * len(q)
        assert r.counters["max_frontier"] > 0
    g.close()
