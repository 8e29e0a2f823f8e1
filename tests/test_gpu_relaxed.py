"""Relaxed matching on the GPU (threshold T > 0): the count of every relaxed
query equals the oracle's count of that edge subset as a query, closed forms
on a cycle, the subset lattice's monotonicity, the GPU filter at T against the
oracle, and the command-line front end's relaxed lines and per-query
isolation."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from gem_amd import Graph, Query, edge_bits, relaxed_queries

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gem_amd", "_lib", "gem")
CAP = 10 ** 7


@pytest.fixture(scope="module")
def er_graph():
    return Graph.erdos_renyi(3000, 15000, 4, 77)


def _queries(g, n, sizes, seed):
    out = []
    s = seed
    while len(out) < n:
        q = Query.random_walk(g, sizes[len(out) % len(sizes)], s)
        s += 1
        if q.num_e >= q.num_v:  # has a cycle, so T > 0 relaxes something
            out.append(q)
    return out


@pytest.mark.parametrize("T,share", [(1, True), (2, True), (1, False), (2, False)])
def test_relaxed_counts_equal_oracle(gpu_ctx, er_graph, T, share):
    """Shared (one filter at T, one table per edge) and one-by-one relaxed
    queries both give the oracle's count of every edge subset."""
    gpu_ctx.load_data(er_graph)
    gpu_ctx.set_relaxed_sharing(share)
    try:
        _check_relaxed(gpu_ctx, er_graph, T)
    finally:
        gpu_ctx.set_relaxed_sharing(True)


def _check_relaxed(gpu_ctx, er_graph, T):
    for q in _queries(er_graph, 4, [5, 6], 300 + T):
        print("relaxed", T, q.edges.tolist(), flush=True)
        res = gpu_ctx.match_relaxed(q, T, cap=CAP)
        assert [r["mask"] for r in res] == relaxed_queries(q, T)
        assert len(res) > 1
        by_mask = {}
        for r in res:
            ref = oracle.count(er_graph, q.subquery(r["mask"]), cap=CAP)
            assert not r["work_limited"]
            assert (r["count"], r["saturated"]) == (ref["count"], ref["saturated"]), (bin(r["mask"]), r, ref)
            by_mask[r["mask"]] = r["count"]
        # leaving out an edge never loses an embedding
        for m, c in by_mask.items():
            for i in range(q.num_e):
                if (m >> i) & 1 and (m & ~(1 << i)) in by_mask:
                    assert by_mask[m & ~(1 << i)] >= c


def test_relaxed_triangle_on_a_cycle(gpu_ctx):
    """Triangle query on the n-cycle at T = 1: no triangle, and each of the
    three 2-edge paths has 2n embeddings (a centre, then its two neighbours in
    either order)."""
    n = 1000
    g = Graph.from_edges(n, [(i, (i + 1) % n) for i in range(n)], np.zeros(n, np.uint32))
    gpu_ctx.load_data(g)
    tri = Query(np.zeros(3), [(0, 1), (0, 2), (1, 2)])
    res = gpu_ctx.match_relaxed(tri, 1)
    assert [(r["mask"], r["count"]) for r in res] == [(0b111, 0), (0b110, 2 * n), (0b101, 2 * n), (0b011, 2 * n)]


def test_relaxed_streamed_tables(gpu_ctx, er_graph):
    """With streamed tables every relaxed query is filtered and tabled on its
    own edges; counts equal the shared form's."""
    gpu_ctx.load_data(er_graph)
    q = _queries(er_graph, 1, [6], 950)[0]
    shared = gpu_ctx.match_relaxed(q, 2, cap=CAP)
    gpu_ctx.set_table_stream(64 << 20, 64 << 10)
    try:
        streamed = gpu_ctx.match_relaxed(q, 2, cap=CAP)
        assert gpu_ctx.stream_stats()["blocks_loaded"] > 0
    finally:
        gpu_ctx.set_table_stream(0)
    assert [(r["mask"], r["count"]) for r in shared] == [(r["mask"], r["count"]) for r in streamed]


def test_relaxed_with_model_and_restarts(gpu_ctx, er_graph):
    """The learned ranking and restarts change no relaxed count."""
    from gem_amd.learned import OrderModel
    gpu_ctx.load_data(er_graph)
    q = _queries(er_graph, 1, [7], 900)[0]
    plain = gpu_ctx.match_relaxed(q, 2, cap=CAP)
    model = OrderModel(gpu_ctx, seed=3)
    try:
        learned = gpu_ctx.match_relaxed(q, 2, model=model, max_orders=6, cap=CAP, first_attempt=1000)
    finally:
        model.close()
    assert [(r["mask"], r["count"], r["saturated"]) for r in plain] == \
           [(r["mask"], r["count"], r["saturated"]) for r in learned]


@pytest.mark.parametrize("T", [1, 2])
def test_filter_at_threshold_equals_oracle(gpu_ctx, er_graph, T):
    gpu_ctx.load_data(er_graph)
    for q in _queries(er_graph, 3, [5, 7], 500):
        num, lists = gpu_ctx.filter(q, threshold=T)
        ref = oracle.filter(er_graph, q, threshold=T)
        for u in range(q.num_v):
            assert num[u] == len(ref[u])
            assert np.array_equal(lists[u], ref[u])


def test_cli_relaxed_lines_and_isolation(tmp_path, er_graph):
    """The front end at -t 1 prints one line per relaxed query with the
    oracle's count; in a directory, a query file it cannot load fails in its
    own child and the other queries still run; a binary CSR cache loads in
    place of the text graph."""
    assert os.path.exists(CLI)
    er_graph.save(tmp_path / "d.graph")
    qd = tmp_path / "q"
    qd.mkdir()
    qs = _queries(er_graph, 2, [5], 700)
    for i, q in enumerate(qs):
        q.save(qd / f"Q_{i}.in")
    (qd / "Q_5.in").write_text("t 2 1\nv 0 0 1\nv 1 0 1\ne 0 7\n")  # endpoint out of range
    print("cli", flush=True)
    r = subprocess.run([CLI, "-q", str(qd), "-d", str(tmp_path / "d.graph"), "-t", "1", "--cap", str(CAP)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "success=2 failed=1" in r.stdout
    assert re.search(r"\[batch-fail\] query 5 file=Q_5\.in exit=\d+ continuing", r.stdout)
    got = re.findall(r"Relaxed Query ET ([01]+): (\d+) matches", r.stdout)
    want = []
    for q in qs:
        for m in relaxed_queries(q, 1):
            want.append((edge_bits(m, q.num_e), str(oracle.count(er_graph, q.subquery(m), cap=CAP)["count"])))
    assert got == want
    # the same data graph from its binary CSR cache, one query at -t 0
    er_graph.save_binary(tmp_path / "d.bin")
    r = subprocess.run([CLI, "-q", str(qd / "Q_0.in"), "-d", str(tmp_path / "d.bin"), "-t", "0", "--cap", str(CAP),
                        "--export", str(tmp_path / "q0.emb")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = re.findall(r"Relaxed Query ET ([01]+): (\d+) matches", r.stdout)
    n0 = oracle.count(er_graph, qs[0], cap=CAP)["count"]
    assert got == [(edge_bits((1 << qs[0].num_e) - 1, qs[0].num_e), str(n0))]
    from gem_amd.export import read_embeddings
    rows, head = read_embeddings(tmp_path / "q0.emb")
    assert head["rows"] == rows.shape[0] == n0 and f"Embeddings written: {n0} " in r.stdout
