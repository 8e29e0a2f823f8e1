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


@pytest.fixture(scope="module")
def hub_graph():
    """30,000 vertices, 3 labels, ~60,000 random edges plus 8 hubs of degree
    3,000-9,000: candidate rows longer than one 2,048-entry table chunk and
    driving lists longer than the walk kernels' 1,024-entry heavy threshold."""
    rng = np.random.default_rng(77)
    n = 30000
    pairs = set()
    while len(pairs) < 60000:
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    for h in range(8):
        for b in rng.choice(np.arange(8, n), size=int(rng.integers(3000, 9000)), replace=False):
            pairs.add((h, int(b)))
    labels = rng.integers(0, 3, n).astype(np.uint32)
    return Graph.from_edges(n, sorted(pairs), labels)


def _hub_queries(g):
    lab = g.labels
    l0, l1, l2 = int(lab[0]), int(lab[1]), int(lab[2])
    return [Query([l0, l1, l2], [(0, 1), (1, 2)]),
            Query([l0, l1, l2], [(0, 1), (0, 2), (1, 2)]),
            Query([l1, l0, l2, l1], [(0, 1), (1, 2), (2, 3), (0, 3)])]


def test_edge_tables_hub_rows(gpu_ctx, hub_graph):
    """Tables whose rows span several chunks of the chunked heavy-row pass."""
    g = hub_graph
    assert int(np.diff(g.offsets).max()) > 2 * 2048
    gpu_ctx.load_data(g)
    for q in _hub_queries(g):
        print("hub tables query", q.labels.tolist(), q.edges.tolist(), flush=True)
        num, lists = gpu_ctx.filter(q)
        ref_can = oracle.filter(g, q)  # hubs of degree > 4096 take the chunked filter path
        for u in range(q.num_v):
            assert np.array_equal(lists[u], ref_can[u])
        sizes = gpu_ctx.build_edge_tables()
        assert np.array_equal(sizes, oracle.table_sizes(g, q))
        deg = np.bincount(q.edges.reshape(-1), minlength=q.num_v)
        for d in range(2 * q.num_e):
            key = q.edges[d // 2][d % 2]
            if deg[key] == 1:
                continue
            off, val = gpu_ctx.read_edge_table(d, int(num[key]), sizes[d])
            ref = oracle.edge_table(g, q, d, int(num[key]), sizes[d])
            assert np.array_equal(off, ref[0])
            assert np.array_equal(val, ref[2])


def test_counts_hub_rows(gpu_ctx, hub_graph):
    """Plain and batched counts on hub rows (heavy walk kernels) equal the oracle's."""
    g = hub_graph
    gpu_ctx.load_data(g)
    qs = _hub_queries(g)
    res = gpu_ctx.match_batch(qs)
    for q, r in zip(qs, res):
        ref = oracle.count(g, q)
        print("hub count query", q.labels.tolist(), q.edges.tolist(), ref["count"], r["count"], flush=True)
        assert r["count"] == ref["count"] and not r["saturated"] and not r["work_limited"]
        c, s, _ = gpu_ctx.match(q)
        assert (c, s) == (ref["count"], False)
