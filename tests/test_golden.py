"""Committed golden fixtures (tests/golden/, written by make_golden.py).

CPU: the fixtures load through the reference-format readers, every stored
embedding is a valid injective label/edge-preserving map (checked here in
numpy, independently of the oracle), their number equals the brute-force
count, and the oracle still reproduces every stored figure.
GPU: the HIP path reproduces counts, GEM order, candidate sizes, edge-table
sizes, per-step rows and the sorted embedding set; the batched path the counts.
"""
import json
import os

import numpy as np
import pytest

import oracle
from gem_amd import Graph, Query
from gem_amd.graph import gem_order

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "expected.json")) as f:
        exp = json.load(f)
    g = Graph.load(os.path.join(HERE, exp["graph"]))
    qs = [Query.load(os.path.join(HERE, e["query"])) for e in exp["queries"]]
    embs = np.load(os.path.join(HERE, "embeddings.npz"))
    return g, qs, exp["queries"], {k: embs[k] for k in embs.files}


def _edge_set(g):
    off, nb = g.offsets, g.neighbors
    src = np.repeat(np.arange(g.num_vertices, dtype=np.uint64), np.diff(off).astype(np.int64))
    return set(zip(src.tolist(), nb[:int(off[-1])].tolist()))


def test_golden_embeddings_valid(golden):
    g, qs, exp, embs = golden
    edges = _edge_set(g)
    for k, (q, e) in enumerate(zip(qs, exp)):
        emb = embs[f"q{k}"]
        assert emb.shape == (e["count"], q.num_v)
        assert e["count"] == e["bruteforce"]
        # sorted, distinct rows
        if len(emb) > 1:
            assert np.all(np.lexsort(emb.T[::-1]) == np.arange(len(emb)))
            assert len(np.unique(emb, axis=0)) == len(emb)
        for row in emb:
            assert len(set(row.tolist())) == q.num_v
            assert all(int(g.labels[row[u]]) == int(q.labels[u]) for u in range(q.num_v))
            assert all((int(row[a]), int(row[b])) in edges for a, b in q.edges)


def test_oracle_reproduces_golden(golden):
    g, qs, exp, embs = golden
    for k, (q, e) in enumerate(zip(qs, exp)):
        r = oracle.count(g, q)
        assert r["count"] == e["count"]
        assert [int(x) for x in r["order"]] == e["gem_order"]
        assert [int(x) for x in r["step_rows"]] == e["step_rows"]
        assert [len(c) for c in oracle.filter(g, q)] == e["num_can"]
        assert [int(x) for x in oracle.table_sizes(g, q)] == e["table_sizes"]
        assert np.array_equal(oracle.enumerate_embeddings(g, q, r["order"]), embs[f"q{k}"])


@pytest.mark.gpu
def test_hip_matches_golden(gpu_ctx, golden):
    g, qs, exp, embs = golden
    gpu_ctx.load_data(g)
    for k, (q, e) in enumerate(zip(qs, exp)):
        print("golden query", e["query"], flush=True)
        num, _ = gpu_ctx.filter(q, want_lists=False)
        assert num.tolist() == e["num_can"]
        sizes = gpu_ctx.build_edge_tables()
        assert sizes.tolist() == e["table_sizes"]
        order = gem_order(q, np.maximum(sizes[0::2], sizes[1::2]))
        assert [int(x) for x in order] == e["gem_order"]
        res = gpu_ctx.join(order, log=True, embeddings_capacity=e["count"] + 1)
        assert res["count"] == e["count"] and not res["saturated"]
        assert res["step_rows"].tolist() == e["step_rows"]
        emb = res["embeddings"]
        got = emb[np.lexsort(emb.T[::-1])]
        assert np.array_equal(got, embs[f"q{k}"])
    counts = [r["count"] for r in gpu_ctx.match_batch(qs)]
    assert counts == [e["count"] for e in exp]
