"""Relaxed queries (threshold T > 0) on the host: the library's edge sets
against the oracle's restatement of the reference's breadth-first relax()
(in-core/lib/plan/src/relax.cpp:84-142), against brute force over edge subsets
and against closed forms; the filter's threshold semantics by definition; the
command-line front end's per-query isolation (no GPU needed: every child then
fails to open one, and the batch still runs to its summary)."""
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle
from gem_amd import GemError, Graph, Query, edge_bits, relaxed_queries

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gem_amd", "_lib", "gem")


def _spans(nv, edges, mask):
    comp = list(range(nv))

    def find(x):
        while comp[x] != x:
            x = comp[x]
        return x

    for i, (a, b) in enumerate(edges):
        if (mask >> i) & 1:
            comp[find(int(a))] = find(int(b))
    return len({find(v) for v in range(nv)}) == 1


def _bruteforce(q, T):
    full = (1 << q.num_e) - 1
    out = []
    for k in range(0, min(T, q.num_e) + 1):
        for drop in itertools.combinations(range(q.num_e), k):
            m = full & ~sum(1 << e for e in drop)
            if _spans(q.num_v, q.edges, m):
                out.append(m)
    return out


def _queries():
    qs = []
    for g in (Graph.erdos_renyi(400, 2400, 3, 5), Graph.erdos_renyi(40, 400, 3, 6)):  # sparse and dense queries
        qs += [Query.random_walk(g, k, s) for k in (4, 5, 6, 7) for s in range(6)]
    return [q for q in qs if q.num_e >= 2]


def test_relaxed_sets_equal_reference_bfs_and_bruteforce():
    n_dense = 0
    for q in _queries():
        n_dense += q.num_e > q.num_v
        for T in range(0, 4):
            lib = relaxed_queries(q, T)
            assert len(set(lib)) == len(lib)
            assert set(lib) == oracle.relax(q, T), (q.edges.tolist(), T)
            # brute force lists by left-out count, then ascending left-out tuple: the library's order
            assert lib == _bruteforce(q, T), (q.edges.tolist(), T)
    assert n_dense >= 5


def test_relaxed_closed_forms():
    K4 = Query(np.zeros(4), [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
    # 1 full; 6 without one edge; all 15 without two (no vertex loses all 3 edges);
    # without three: the 16 spanning trees of K4 (Cayley 4^2), not the 4 triangles
    assert [len(relaxed_queries(K4, T)) for T in range(6)] == [1, 7, 22, 38, 38, 38]
    for n in range(3, 9):
        cyc = Query(np.zeros(n), [(i, (i + 1) % n) if i + 1 < n else (0, n - 1) for i in range(n)])
        assert [len(relaxed_queries(cyc, T)) for T in range(4)] == [1, n + 1, n + 1, n + 1]
        path = Query(np.zeros(n), [(i, i + 1) for i in range(n - 1)])
        assert relaxed_queries(path, 3) == [(1 << (n - 1)) - 1]  # a tree has nothing to leave out
    tri = Query(np.zeros(3), [(0, 1), (0, 2), (1, 2)])
    assert relaxed_queries(tri, 1) == [0b111, 0b110, 0b101, 0b011]


def _kirchhoff(q):
    L = np.zeros((q.num_v, q.num_v))
    for a, b in q.edges:
        L[a, a] += 1
        L[b, b] += 1
        L[a, b] -= 1
        L[b, a] -= 1
    return int(round(np.linalg.det(L[1:, 1:])))


def test_spanning_tree_level_is_the_matrix_tree_count():
    """Leaving out num_e - num_v + 1 edges leaves exactly the spanning trees,
    whose number is the Laplacian's cofactor (Kirchhoff)."""
    seen = 0
    for q in _queries():
        k = q.num_e - q.num_v + 1
        if k < 1 or k > 5:
            continue
        sets = relaxed_queries(q, k)
        trees = [m for m in sets if bin(m).count("1") == q.num_v - 1]
        assert len(trees) == _kirchhoff(q)
        seen += 1
    assert seen >= 10


def test_relaxed_limits_and_bad_queries():
    K6 = Query(np.zeros(6), [(i, j) for i in range(6) for j in range(i + 1, 6)])
    n = len(relaxed_queries(K6, 4))
    assert n == sum(1 for _ in _bruteforce(K6, 4))
    lib = __import__("gem_amd").load()
    import ctypes as C
    v = K6.view()
    got = C.c_uint32()
    masks = (C.c_uint64 * (4 * 8))()
    assert lib.gem_plan_relaxed_queries(C.byref(v), 4, 8, masks, C.byref(got)) != 0  # more than 8
    one_edge = Query(np.zeros(2), [(0, 1)])
    with pytest.raises(GemError):
        relaxed_queries(one_edge, 1)


def test_edge_bits_print_like_a_32_bit_bitset():
    assert edge_bits(0b101, 3) == "0" * 29 + "101"
    assert len(edge_bits((1 << 40) - 1, 40)) == 40


@pytest.mark.parametrize("T", [1, 2])
def test_filter_threshold_by_definition(T):
    """Can(u) at threshold T: same label, deg(v) + T >= deg(u), and the
    neighbour-label deficit sum_l max(0, nlc_u[l] - nlc_v[l]) <= T
    (in-core/lib/filter/src/filter_kernel.cu:205, 255)."""
    g = Graph.erdos_renyi(300, 1200, 3, 21)
    qs = [Query.random_walk(g, k, 900 + k) for k in (4, 5, 6)]
    for q in qs:
        can = oracle.filter(g, q, threshold=T)
        can0 = oracle.filter(g, q)
        qdeg = np.bincount(q.edges.reshape(-1), minlength=q.num_v)
        qnlc = np.zeros((q.num_v, 256), int)
        for s, d in q.edges:
            qnlc[s, q.labels[d]] += 1
            qnlc[d, q.labels[s]] += 1
        qlabels = set(q.labels.tolist())
        grew = 0
        for u in range(q.num_v):
            want = []
            for v in range(g.num_vertices):
                if g.labels[v] != q.labels[u]:
                    continue
                nb = g.neighbors[g.offsets[v]:g.offsets[v + 1]]
                if len(nb) + T < qdeg[u]:
                    continue
                dn = np.zeros(256, int)
                for w in nb:
                    if int(g.labels[w]) in qlabels:
                        dn[g.labels[w]] += 1
                if np.maximum(qnlc[u] - dn, 0).sum() <= T:
                    want.append(v)
            assert can[u].tolist() == want
            assert set(can0[u].tolist()) <= set(want)
            grew += len(want) > len(can0[u])
        assert grew >= 1


def test_threshold_filter_is_sound_for_every_relaxed_query():
    """Every embedding of every relaxed query at T uses, for each query vertex,
    a candidate of the query's own filter at threshold T: one filter serves
    the whole lattice."""
    g = Graph.erdos_renyi(80, 320, 2, 31)
    for q in [Query.random_walk(g, 5, s) for s in range(40, 46)]:
        for T in (1, 2):
            can = [set(c.tolist()) for c in oracle.filter(g, q, threshold=T)]
            for m in relaxed_queries(q, T):
                rq = q.subquery(m)
                emb = oracle.enumerate_embeddings(g, rq, oracle.count(g, rq)["order"])
                for u in range(q.num_v):
                    assert set(emb[:, u].tolist()) <= can[u]


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except Exception:
        return True


@pytest.mark.skipif(not os.path.exists(CLI), reason="command-line front end not built")
def test_cli_directory_isolates_each_query(tmp_path):
    """Each query of a directory runs in a child; a child that fails is
    reported and the batch continues to its summary (exit 0), as the
    reference's directory runner does.  Without a GPU every child fails."""
    if not _no_gpu():
        pytest.skip("a GPU is present: covered by the GPU tests")
    g = Graph.erdos_renyi(200, 800, 3, 1)
    g.save(tmp_path / "d.graph")
    qd = tmp_path / "q"
    qd.mkdir()
    for i in (7, 2):
        Query.random_walk(g, 5, i).save(qd / f"Q_{i}.in")
    r = subprocess.run([CLI, "-q", str(qd), "-d", str(tmp_path / "d.graph"), "-t", "1"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "2" and "[batch-fail] query 2 file=Q_2.in exit=1 continuing" in lines
    assert "[batch-fail] query 7 file=Q_7.in exit=1 continuing" in lines
    assert "success=0 failed=2" in lines
    (qd / "notes.in").write_text("x")
    r = subprocess.run([CLI, "-q", str(qd), "-d", str(tmp_path / "d.graph")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "does not match expected pattern" in r.stderr
    r = subprocess.run([CLI, "-q", str(qd / "Q_2.in"), "-d", str(tmp_path / "d.graph"), "-t", "11"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
