"""Host-resident data graph: the neighbour array stays in pinned host memory
and the filter and table kernels read it over PCIe.  Candidates, tables and
counts equal the oracle's, alone and together with edge tables streamed
through a small device cache (the out-of-core form: graph and tables on the
host)."""
import numpy as np
import pytest

import oracle
from gem_amd import Graph, Query

pytestmark = pytest.mark.gpu


def _queries(g, n, sizes, seed):
    return [Query.random_walk(g, sizes[i % len(sizes)], seed + i) for i in range(n)]


@pytest.fixture
def host_ctx(gpu_ctx):
    gpu_ctx.set_graph_residency(True)
    try:
        yield gpu_ctx
    finally:
        gpu_ctx.set_graph_residency(False)
        gpu_ctx.set_table_stream(0)


def test_host_graph_filter_tables_counts(host_ctx):
    g = Graph.rmat(16, 16.0, 8, 4242)
    host_ctx.load_data(g)
    for q in _queries(g, 4, [5, 6, 8], 31):
        print("host graph query", q.num_v, q.edges.tolist(), flush=True)
        num, lists = host_ctx.filter(q)
        ref = oracle.filter(g, q)
        for u in range(q.num_v):
            assert np.array_equal(lists[u], ref[u])
        sizes = host_ctx.build_edge_tables()
        assert np.array_equal(sizes, oracle.table_sizes(g, q))
        r = host_ctx.match_full(q, cap=10 ** 8)
        want = oracle.count(g, q, cap=10 ** 8)
        assert (r["count"], r["saturated"]) == (want["count"], want["saturated"])


def test_host_graph_with_streamed_tables_batched(host_ctx):
    """Graph on the host and tables streamed through a 128 MiB cache of 64 KiB
    slots (2048 slots, at least 256 rows x 7 backward edges, the most an
    8-vertex query has): the batched counts and per-level rows equal the
    oracle's walk."""
    g = Graph.rmat(17, 16.0, 16, 99)
    host_ctx.load_data(g)
    qs = _queries(g, 10, [4, 6, 8], 77)
    capv = 10 ** 8
    vos = [oracle.vertex_order_from_edges(q, oracle.count(g, q, cap=capv)["order"]) for q in qs]
    host_ctx.set_table_stream(128 << 20, 64 << 10)
    got = host_ctx.match_batch(qs, vertex_orders=vos, cap=capv, level_budget=64 << 20)
    st = host_ctx.stream_stats()
    assert st["blocks_loaded"] > 0
    for q, vo, r in zip(qs, vos, got):
        ref = oracle.vcount(g, q, vo, cap=capv)
        assert (r["count"], r["saturated"]) == (ref["count"], ref["saturated"])
        if not ref["saturated"]:
            assert [int(x) for x in r["level_rows"]] == [int(x) for x in ref["level_rows"]]
