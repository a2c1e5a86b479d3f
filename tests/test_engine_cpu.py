"""The whole engine (device logic + host engine) on the CPU build of libgst.

No GPU is needed: tests/cpu_sim compiles the HIP device code and the host
engine for the host (wave lanes as threads, host memory as device memory).
This pins the engine's logic -- group-distance jobs (multi-label level 1,
STAR, shortest-path reconstruction from C), the hop-bounded per-query level 1
and its pruning (bounds per state height, hop star and path upper bounds),
batching and admission over several jobs -- against the golden fixtures and
the oracle.  The GPU parity tests check the same properties on the MI355X.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))


@pytest.fixture(scope="module")
def cpu_results():
    import cpu_sim
    try:
        lib = cpu_sim.lib_path()
    except (OSError, subprocess.CalledProcessError) as e:  # no g++
        pytest.skip(f"CPU build unavailable: {e}")
    env = dict(os.environ, GST_LIB=lib)
    # every 30th tiny case (all D; the GPU tests take all of them), 54 k=5
    # queries = 270 labels = two jobs, and 16 hop-bounded queries; the checks
    # run in three processes side by side (each emulated wave is 64 threads)
    groups = ["tiny,hop,branching,static", "jobs,robust,graph", "cache,progressive,backbone"]
    procs = [subprocess.Popen([sys.executable, os.path.join(ROOT, "tests", "cpu_sim", "run_checks.py"),
                               "30", "54", "16", grp], env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                              text=True) for grp in groups]
    out = {}
    for p in procs:
        so, se = p.communicate(timeout=900)
        assert p.returncode == 0, se[-2000:]
        out.update(json.loads(so.strip().splitlines()[-1]))
    return out


def test_tiny_cases_on_cpu_build(cpu_results):
    t = cpu_results["tiny"]
    assert t["checked"] > 30
    assert t["bad"] == [], t["bad"]


def test_group_distance_jobs_on_cpu_build(cpu_results):
    j = cpu_results["jobs"]
    assert j["labels"] > 256  # more than one job
    assert j["bad"] == [], j["bad"]
    # source chunks: the dirty 16-label chunks of the frontier entries, each
    # relaxed along every arc of its entry (16 labels per relaxation) and at
    # most 16 chunks (256 labels) per entry
    c = j["counters"]
    assert 0 < 16 * c["src_chunks"] <= c["relaxations"], c
    assert c["src_chunks"] <= 16 * c["entries"], c
    # this graph's hub distances fit 16 bits: the fill reads 2 bytes per hub,
    # 4 with the packed fill switched off
    assert (c["hub_bytes"], c["hub_bytes_32"]) == (2, 4), c


def test_hop_queries_on_cpu_build(cpu_results):
    h = cpu_results["hop"]
    assert h["feasible"] >= h["checked"] // 2
    assert h["bad"] == [], h["bad"]


def test_grown_branch_vertex_on_cpu_build(cpu_results):
    b = cpu_results["branching"]
    assert b["ok"] and b["cost"] == b["expect"] == 36, b


def test_graph_replay_on_cpu_build(cpu_results):
    r = cpu_results["graph"]
    assert r["eager"] == r["expect"]
    assert all(x == r["expect"] for x in r["graph"]), r


def test_static_entry_on_cpu_build(cpu_results):
    r = cpu_results["static"]
    assert r["checked"] == 5 and r["bad"] == [], r
    assert r["small_buffer_error"], r


def test_failure_handling_on_cpu_build(cpu_results):
    r = cpu_results["robust"]
    assert r["mix_ok"], r
    assert r["count_ok"], r
    assert r["deadline_ok"], r
    assert r["poison_ok"], r
    assert r["server_ok"], r
    assert r["server_stats"][0] >= r["server_stats"][1] > 0, r
    assert r["regroup_server_ok"], r


def test_group_distance_cache_on_cpu_build(cpu_results):
    c = cpu_results["cache"]
    assert c["exact"] and c["single_exact"] and c["regroup_exact"], c
    assert c["hits_first"] == 0, c  # columns become readable only when their job ends
    assert c["inserts"] > 0 and c["hits_later"] > 0 and c["per_query"] == c["hits_later"], c


def test_progressive_answers_on_cpu_build(cpu_results):
    p = cpu_results["progressive"]
    assert p["bad"] == [], p


def test_backbone_on_cpu_build(cpu_results):
    r = cpu_results["backbone"]
    assert r["graphs"] == 4 and r["queries"] == 10, r
    assert r["removed"] > 0, r
    assert r["bad"] == [], r["bad"]
