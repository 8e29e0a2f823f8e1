"""CPU tests of the batched join's host scheduler and the caching allocator.

tests/native/sched_sim.cpp compiles the library's own scheduler
(gem_amd/csrc/host/vsched.h) and allocator bookkeeping
(gem_amd/csrc/host/pool_core.h) against a simulated device: levels exist only
as row counts, 64-row units get hash-drawn output bounds with rare hub units
~10^4 times larger, and buffers are fake addresses from a PoolCore.  No GPU.
"""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "gem_amd", "csrc", "host")
SRC = os.path.join(ROOT, "tests", "native", "sched_sim.cpp")


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sched") / "sched_sim")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", HOST, SRC, "-o", exe], check=True)
    return exe


def _run(exe, *args):
    out = subprocess.run([exe, *map(str, args)], check=True, capture_output=True, text=True, timeout=300).stdout
    return [json.loads(line) for line in out.splitlines() if line.strip()]


def test_plan_chunks_cases(sim):
    (r,) = _run(sim, "plan")
    # per-query quota split over the levels still to write; a unit larger
    # than the whole free quota waits
    assert r["share"] == {"forced": -1, "end": [2, 0]}
    # no budget, no count-only work: the deepest level is forced one unit
    assert r["forced"] == {"forced": 1, "end": [0, 1]}
    # no budget but count-only work runs: everyone waits
    assert r["wait"] == {"forced": -1, "end": [0, 0]}
    # units with a zero output bound need no budget
    assert r["zero"] == {"forced": -1, "end": [2]}
    # a forced chunk extends up to force_chunk bytes, not into the hub unit
    assert r["force_chunk"] == {"forced": 1, "end": [0, 3]}
    # one query's chunk never exceeds max_chunk
    assert r["max_chunk"] == {"forced": -1, "end": [2]}
    # a share below one unit takes one unit when the free quota holds it
    assert r["one_unit"] == {"forced": -1, "end": [1]}
    # pending rows are charged to the quota
    assert r["quota_live"] == {"forced": -1, "end": [1]}


def test_pool_class_boundary(sim):
    """Requests just below 1 MiB used to get a small-class size of exactly
    1 MiB (or a carved size below it) and were then freed down the other path,
    never to be reused: every such buffer leaked a driver allocation."""
    (r,) = _run(sim, "pool")
    for rnd in r["rounds"]:
        assert rnd["classes_ok"]
        assert rnd["in_use_full"] == rnd["expect"]
        assert rnd["in_use_after"] == 0
    # retaining trims: held drops to at most the retained amount, by whole
    # segments and blocks, and a second such trim releases nothing
    assert r["held_full"] > (10 << 30) >= r["held_retain"] > 0
    assert r["released_retain"] == r["held_full"] - r["held_retain"] and r["released_again"] == 0
    assert r["held_after_trim"] == 0


def test_scheduler_memory_bounded_over_passes(sim):
    """160 queries of 4-32 vertices with hub-heavy units, a 256 MiB level
    budget, three passes on one pool: the charged bytes stay within
    budget + (levels - 1) x max(force chunk, largest forced unit), the pool's
    peak within its rounding of that, and the pool returns to its starting
    state after every pass."""
    budget = 256 << 20
    passes = _run(sim, "run", 160, budget, 3, 11)
    assert len(passes) == 3
    for p in passes:
        headroom = (p["max_nv"] - 1) * max(p["force_chunk"], p["max_unit_bytes"])
        assert p["peak_charged"] <= budget + headroom
        assert p["max_overshoot"] <= headroom
        # rounding: 64 KiB units for segment blocks, size classes below 1 MiB
        assert p["pool_peak"] <= 1.25 * p["peak_charged"] + (1 << 20)
        assert p["pool_in_use_after"] == 0
        assert p["held_after_trim"] == 0
        assert p["pool_in_use_before"] == passes[0]["pool_in_use_before"]
        assert p["pending"] == 0
        assert p["saturated"] + p["work_limited"] + p["finished"] == p["queries"]
        # the scenario exercises waiting and the forced path
        assert p["waits"] > 0 and p["forced"] > 0


def test_scheduler_large_budget_never_forces(sim):
    """With a budget far above what the levels need, nothing waits past it."""
    (p,) = _run(sim, "run", 48, 64 << 30, 1, 5)
    assert p["forced"] == 0
    assert p["max_overshoot"] == 0
    assert p["pool_in_use_after"] == 0 and p["held_after_trim"] == 0


def test_scheduler_continuous_admission(sim):
    """160 queries, at most 16 running at once (the others start as running
    ones end): every query ends, the charged bytes stay within the same
    bound, and the pool returns to its starting state."""
    budget = 256 << 20
    passes = _run(sim, "run", 160, budget, 2, 13, 16)
    for p in passes:
        headroom = (p["max_nv"] - 1) * max(p["force_chunk"], p["max_unit_bytes"])
        assert p["peak_charged"] <= budget + headroom
        assert p["pending"] == 0
        assert p["saturated"] + p["work_limited"] + p["finished"] == p["queries"]
        assert p["pool_in_use_after"] == 0 and p["held_after_trim"] == 0


def test_scheduler_swaps_levels_to_host(sim):
    """Swapping on (1 GiB of host memory, a 64 MiB device budget, 160 queries,
    16 running): levels move out and back, the host bytes stay within their
    bound and are all released, every query ends, no level is measured or
    extended while in host memory (the simulator exits if one is), a level
    comes back with its own rows (size checked), and the device-resident peak
    is no higher than without swapping."""
    budget, swap = 64 << 20, 1 << 30
    (p,) = _run(sim, "run", 160, budget, 1, 11, 16, swap)
    (q,) = _run(sim, "run", 160, budget, 1, 11, 16, 0)
    assert p["swaps_out"] > 0 and p["swaps_in"] > 0 and q["swaps_out"] == 0
    assert 0 < p["peak_host"] <= swap and p["backend_host_peak"] == p["peak_host"]
    assert p["host_left"] == 0
    assert p["pending"] == 0 and p["saturated"] + p["work_limited"] + p["finished"] == p["queries"]
    assert p["pool_in_use_after"] == 0
    assert p["peak_charged"] <= q["peak_charged"]


def test_scheduler_lazy_admission(sim):
    """160 queries, 16 running, the first 24 ready at the start and the rest
    prepared elsewhere one per 5,000 extend steps: no level of a query is
    measured, extended or counted before it is ready (the simulator exits if
    one is), the schedule skips a query that is not ready while others run and
    waits for it only when nothing runs (it exits otherwise), every query
    starts and ends, and the charged bytes keep their bound."""
    budget = 256 << 20
    (p,) = _run(sim, "run", 160, budget, 1, 13, 16, 0, 24, 5000)
    assert p["not_ready"] > 0 and p["ready_waits"] > 0
    assert p["started"] == p["queries"] == 160
    assert p["pending"] == 0 and p["saturated"] + p["work_limited"] + p["finished"] == p["queries"]
    headroom = (p["max_nv"] - 1) * max(p["force_chunk"], p["max_unit_bytes"])
    assert p["peak_charged"] <= budget + headroom
    assert p["pool_in_use_after"] == 0 and p["held_after_trim"] == 0
