"""Pins for the CPU oracle: closed-form counts, brute force, and the filter /
table definitions.  The oracle decides GPU parity, so each of its
functions is checked against something it does not compute itself."""
import itertools
import math

import numpy as np
import pytest

import oracle
from gem_amd import Graph, Query


def complete_graph(n, labels=None):
    edges = [(i, j) for i in range(n) for j in range(i + 1, n)]
    return Graph.from_edges(n, edges, np.zeros(n, np.uint32) if labels is None else labels)


def cycle_graph(n):
    return Graph.from_edges(n, [(i, (i + 1) % n) for i in range(n)], np.zeros(n, np.uint32))


def grid_graph(a, b):
    idx = lambda i, j: i * b + j
    edges = []
    for i in range(a):
        for j in range(b):
            if i + 1 < a:
                edges.append((idx(i, j), idx(i + 1, j)))
            if j + 1 < b:
                edges.append((idx(i, j), idx(i, j + 1)))
    return Graph.from_edges(a * b, edges, np.zeros(a * b, np.uint32))


def star_graph(n):
    return Graph.from_edges(n + 1, [(0, i) for i in range(1, n + 1)], np.zeros(n + 1, np.uint32))


def path_query(k, lab=0):
    return Query(np.full(k, lab, np.uint32), np.array([(i, i + 1) for i in range(k - 1)], np.uint32))


def clique_query(k):
    return Query(np.zeros(k, np.uint32), np.array([(i, j) for i in range(k) for j in range(i + 1, k)], np.uint32))


def falling(n, k):
    return math.perm(n, k)


@pytest.mark.parametrize("n,k", [(5, 3), (6, 4), (7, 3), (8, 5)])
def test_complete_graph_clique_and_path(n, k):
    g = complete_graph(n)
    assert oracle.count(g, clique_query(k))["count"] == falling(n, k)
    assert oracle.count(g, path_query(k))["count"] == falling(n, k)


@pytest.mark.parametrize("n,k", [(7, 3), (9, 4), (12, 6)])
def test_cycle_path(n, k):
    # each of the 2n directed starts gives one path
    assert oracle.count(cycle_graph(n), path_query(k))["count"] == 2 * n


@pytest.mark.parametrize("a,b", [(3, 3), (4, 5), (6, 7)])
def test_grid_four_cycle(a, b):
    q = Query(np.zeros(4, np.uint32), np.array([(0, 1), (1, 2), (2, 3), (0, 3)], np.uint32))
    assert oracle.count(grid_graph(a, b), q)["count"] == 8 * (a - 1) * (b - 1)


@pytest.mark.parametrize("n,c", [(5, 2), (6, 3), (7, 4)])
def test_star(n, c):
    q = Query(np.zeros(c + 1, np.uint32), np.array([(0, i) for i in range(1, c + 1)], np.uint32))
    assert oracle.count(star_graph(n), q)["count"] == falling(n, c)


@pytest.mark.parametrize("a,b", [(3, 4), (5, 2)])
def test_labelled_bipartite(a, b):
    # K_{a,b} with side labels 0/1; query A-B-A path: a * b * (a-1)
    n = a + b
    edges = [(i, a + j) for i in range(a) for j in range(b)]
    lab = np.array([0] * a + [1] * b, np.uint32)
    g = Graph.from_edges(n, edges, lab)
    q = Query(np.array([0, 1, 0], np.uint32), np.array([(0, 1), (1, 2)], np.uint32))
    assert oracle.count(g, q)["count"] == a * b * (a - 1)
    q2 = Query(np.array([0, 1, 1], np.uint32), np.array([(0, 1), (0, 2)], np.uint32))
    assert oracle.count(g, q2)["count"] == a * b * (b - 1)


def test_triangles_by_matrix():
    g = Graph.erdos_renyi(80, 500, 1, 5)
    A = np.zeros((80, 80), np.int64)
    for v in range(80):
        for w in g.neighbors[g.offsets[v]:g.offsets[v + 1]]:
            A[v, w] = 1
    tri = int(np.trace(A @ A @ A)) // 6
    assert oracle.count(g, clique_query(3))["count"] == 6 * tri


def _small_queries(g, n, ks, seed):
    out = []
    for i in range(n):
        q = Query.random_walk(g, ks[i % len(ks)], seed + i)
        if q.num_e >= 2:
            out.append(q)
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_count_vs_bruteforce(seed):
    g = Graph.erdos_renyi(50, 180, 3, seed)
    for q in _small_queries(g, 15, [3, 4, 5, 6], 100 * seed):
        assert oracle.count(g, q)["count"] == oracle.bruteforce(g, q), (q.labels, q.edges)


def test_enumerate_matches_count_and_is_valid():
    g = Graph.erdos_renyi(60, 220, 2, 9)
    for q in _small_queries(g, 8, [4, 5], 77):
        r = oracle.count(g, q)
        emb = oracle.enumerate_embeddings(g, q, r["order"])
        assert len(emb) == r["count"]
        assert len({tuple(x) for x in emb}) == len(emb)
        for row in emb[:200]:
            assert len(set(row.tolist())) == q.num_v
            for u in range(q.num_v):
                assert g.labels[row[u]] == q.labels[u]
            for s, d in q.edges:
                assert row[d] in g.neighbors[g.offsets[row[s]]:g.offsets[row[s] + 1]]


def test_order_invariance():
    """The final count is the same under every connected edge order."""
    g = Graph.erdos_renyi(60, 260, 2, 4)
    rng = np.random.default_rng(0)
    for q in _small_queries(g, 6, [5, 6], 300):
        base = oracle.count(g, q)["count"]
        for _ in range(4):
            order, used = [], set()
            first = int(rng.integers(q.num_e))
            order.append(first)
            used.add(first)
            verts = set(q.edges[first].tolist())
            while len(order) < q.num_e:
                cand = [e for e in range(q.num_e) if e not in used and (set(q.edges[e].tolist()) & verts)]
                e = int(rng.choice(cand))
                order.append(e)
                used.add(e)
                verts |= set(q.edges[e].tolist())
            assert oracle.count(g, q, order=np.array(order, np.uint32))["count"] == base


def test_step_rows_are_partial_embedding_counts():
    """rows after step s = number of injective, candidate-respecting maps of
    the first s+1 edges (checked by brute force over the prefix subgraph)."""
    g = Graph.erdos_renyi(40, 140, 2, 12)
    for q in _small_queries(g, 5, [4, 5], 500):
        r = oracle.count(g, q)
        can = oracle.filter(g, q)
        order = r["order"]
        for s in range(q.num_e):
            es = [tuple(q.edges[e]) for e in order[:s + 1]]
            vs = sorted({v for e in es for v in e})
            cnt = 0
            for img in itertools.permutations(range(g.num_vertices), len(vs)):
                m = dict(zip(vs, img))
                if any(m[u] not in set(can[u].tolist()) for u in vs):
                    continue
                if all(m[b] in g.neighbors[g.offsets[m[a]]:g.offsets[m[a] + 1]] for a, b in es):
                    cnt += 1
            assert cnt == r["step_rows"][s], (s, q.edges, order)
        break  # one query: the permutation brute force is expensive


def test_filter_by_definition():
    g = Graph.erdos_renyi(300, 1200, 3, 21)
    for q in _small_queries(g, 6, [4, 6], 900):
        can = oracle.filter(g, q)
        qdeg = np.zeros(q.num_v, int)
        qnlc = np.zeros((q.num_v, 256), int)
        for s, d in q.edges:
            qdeg[s] += 1
            qdeg[d] += 1
            qnlc[s, q.labels[d]] += 1
            qnlc[d, q.labels[s]] += 1
        qlabels = set(q.labels.tolist())
        for u in range(q.num_v):
            want = []
            for v in range(g.num_vertices):
                if g.labels[v] != q.labels[u]:
                    continue
                nb = g.neighbors[g.offsets[v]:g.offsets[v + 1]]
                if len(nb) < qdeg[u]:
                    continue
                dn = np.zeros(256, int)
                for w in nb:
                    if int(g.labels[w]) in qlabels:
                        dn[g.labels[w]] += 1
                if np.maximum(qnlc[u] - dn, 0).sum() == 0:
                    want.append(v)
            assert can[u].tolist() == want


def test_filter_is_sound():
    """Every vertex used by an embedding is a candidate of its query vertex."""
    g = Graph.erdos_renyi(60, 240, 2, 31)
    for q in _small_queries(g, 5, [4, 5], 40):
        can = oracle.filter(g, q)
        emb = oracle.enumerate_embeddings(g, q, oracle.count(g, q)["order"])
        for u in range(q.num_v):
            assert set(emb[:, u].tolist()) <= set(can[u].tolist())


def test_tables_by_definition_and_symmetry():
    g = Graph.erdos_renyi(200, 900, 3, 41)
    for q in _small_queries(g, 4, [4, 5], 60):
        can = oracle.filter(g, q)
        sizes = oracle.table_sizes(g, q)
        deg = np.bincount(q.edges.reshape(-1), minlength=q.num_v)
        for i, (s, d) in enumerate(q.edges):
            if deg[s] != 1 and deg[d] != 1:
                assert sizes[2 * i] == sizes[2 * i + 1]
            for dd, (u, un) in ((2 * i, (s, d)), (2 * i + 1, (d, s))):
                if deg[u] == 1:
                    assert sizes[dd] == 0
                    continue
                off, ids, pos = oracle.edge_table(g, q, dd, len(can[u]), sizes[dd])
                cu = set(can[un].tolist())
                want = []
                wo = [0]
                for v in can[u]:
                    row = [w for w in g.neighbors[g.offsets[v]:g.offsets[v + 1]] if w in cu]
                    want += row
                    wo.append(len(want))
                assert ids.tolist() == want and off.tolist() == wo
                assert pos.tolist() == [int(np.searchsorted(can[un], w)) for w in want]
