"""GPU parity of the plain path against the CPU oracle (bit-exact).

Covers the filter (candidate lists), the position index (= lower_bound in
Can(u)), the edge tables (offsets and values), the join (final count, per-step
row counts, sorted embedding sets) and chunked levels, on Erdos-Renyi graphs
shaped like the 10k-vertex / 50k-edge / 8-label workload and on a small R-MAT.
"""
import numpy as np
import pytest

import oracle
from gem_amd import NOT_FOUND, Graph, Query, gem_order

pytestmark = pytest.mark.gpu


def _edge_sizes(sizes):
    return np.maximum(sizes[0::2], sizes[1::2])


@pytest.fixture(scope="module")
def er_graph():
    return Graph.erdos_renyi(10000, 50000, 8, 1001)


@pytest.fixture(scope="module")
def rmat_graph():
    return Graph.rmat(14, 16.0, 8, 1002)


def _queries(g, n, ks, seed):
    out = []
    for i in range(n):
        k = ks[i % len(ks)]
        q = Query.random_walk(g, k, seed + i)
        if q.num_e >= 2:
            out.append(q)
    return out


def test_filter_and_index(gpu_ctx, er_graph):
    gpu_ctx.load_data(er_graph)
    for q in _queries(er_graph, 6, [5, 6, 8], 11):
        print("filter query", q.labels.tolist(), q.edges.tolist(), flush=True)
        num, lists = gpu_ctx.filter(q)
        ref = oracle.filter(er_graph, q)
        for u in range(q.num_v):
            assert num[u] == len(ref[u])
            assert np.array_equal(lists[u], ref[u])
        # position lookup = lower_bound in Can(u) for members, NOT_FOUND otherwise
        u = 0
        probe = np.concatenate([ref[u][:50], np.arange(0, er_graph.num_vertices, 97, dtype=np.uint32)])
        pos = gpu_ctx.position_lookup(u, probe)
        for v, p in zip(probe, pos):
            i = np.searchsorted(ref[u], v)
            want = i if (i < len(ref[u]) and ref[u][i] == v) else NOT_FOUND
            assert p == want


def test_edge_tables(gpu_ctx, er_graph):
    gpu_ctx.load_data(er_graph)
    for q in _queries(er_graph, 4, [5, 7], 21):
        print("tables query", q.labels.tolist(), q.edges.tolist(), flush=True)
        num, lists = gpu_ctx.filter(q)
        sizes = gpu_ctx.build_edge_tables()
        assert np.array_equal(sizes, oracle.table_sizes(er_graph, q))
        deg = np.zeros(q.num_v, int)
        for s, d in q.edges:
            deg[s] += 1
            deg[d] += 1
        for d in range(2 * q.num_e):
            key = q.edges[d // 2][d % 2]
            if deg[key] == 1:
                continue
            off, val = gpu_ctx.read_edge_table(d, int(num[key]), sizes[d])
            ref = oracle.edge_table(er_graph, q, d, int(num[key]), sizes[d])
            assert ref is not None
            assert np.array_equal(off, ref[0])
            assert np.array_equal(val, ref[2])


def test_join_counts_er(gpu_ctx, er_graph):
    gpu_ctx.load_data(er_graph)
    for q in _queries(er_graph, 20, [5], 1001):
        print("join query", q.labels.tolist(), q.edges.tolist(), flush=True)
        count, sat, order = gpu_ctx.match(q)
        ref = oracle.count(er_graph, q)
        assert list(order) == list(ref["order"])
        assert count == ref["count"] and not sat


def test_join_step_rows_and_embeddings(gpu_ctx, er_graph):
    gpu_ctx.load_data(er_graph)
    for q in _queries(er_graph, 6, [4, 5, 6], 3001):
        print("embedding query", q.labels.tolist(), q.edges.tolist(), flush=True)
        gpu_ctx.filter(q, want_lists=False)
        sizes = gpu_ctx.build_edge_tables()
        order = gem_order(q, _edge_sizes(sizes))
        ref = oracle.count(er_graph, q, order=order)
        res = gpu_ctx.join(order, log=True)
        assert res["count"] == ref["count"]
        assert np.array_equal(res["step_rows"], ref["step_rows"])
        if ref["count"] <= 200000:
            emb = gpu_ctx.join(order, embeddings_capacity=ref["count"] + 1)["embeddings"]
            want = oracle.enumerate_embeddings(er_graph, q, order)
            got = emb[np.lexsort(emb.T[::-1])] if emb is not None else np.zeros((0, q.num_v), np.uint32)
            assert np.array_equal(got, want)


def test_join_chunked_levels(gpu_ctx, rmat_graph):
    """A 1 MiB level budget forces every large level to be split."""
    gpu_ctx.load_data(rmat_graph)
    try:
        for q in _queries(rmat_graph, 6, [5, 6], 5001):
            print("chunk query", q.labels.tolist(), q.edges.tolist(), flush=True)
            ref = oracle.count(rmat_graph, q, cap=10 ** 8, work_limit=10 ** 9)
            if ref is None:
                continue
            gpu_ctx.set_level_budget(2 ** 31)
            c_big, s_big, _ = gpu_ctx.match(q, cap=10 ** 8)
            gpu_ctx.set_level_budget(2 ** 20)
            c_small, s_small, _ = gpu_ctx.match(q, cap=10 ** 8)
            assert (c_big, s_big) == (ref["count"], ref["saturated"])
            if not ref["saturated"]:
                assert (c_small, s_small) == (c_big, s_big)
    finally:
        gpu_ctx.set_level_budget(2 ** 31)


def test_cap_semantics(gpu_ctx, rmat_graph):
    gpu_ctx.load_data(rmat_graph)
    q = Query(np.array([0, 0, 0], np.uint32), np.array([[0, 1], [1, 2]], np.uint32))
    ref = oracle.count(rmat_graph, q)
    full, sat, _ = gpu_ctx.match(q)
    assert full == ref["count"] and not sat
    cap = full // 2
    c, s, _ = gpu_ctx.match(q, cap=cap)
    assert s and c == cap
    c, s, _ = gpu_ctx.match(q, cap=full)
    assert not s and c == full


def test_query_errors(gpu_ctx, er_graph):
    from gem_amd import GemError
    gpu_ctx.load_data(er_graph)
    with pytest.raises(GemError):  # single edge: the reference computes nothing
        gpu_ctx.match(Query(np.array([0, 1], np.uint32), np.array([[0, 1]], np.uint32)))
    with pytest.raises(GemError):  # disconnected
        gpu_ctx.match(Query(np.array([0, 1, 2, 3], np.uint32), np.array([[0, 1], [2, 3]], np.uint32)))
    # a label absent from the data graph: no candidates, count 0
    c, s, _ = gpu_ctx.match(Query(np.array([0, 1, 999], np.uint32), np.array([[0, 1], [1, 2]], np.uint32)))
    assert c == 0 and not s
