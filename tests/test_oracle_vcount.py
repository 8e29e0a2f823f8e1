"""Pins for the oracle's vertex-level walk (orc_vcount), against which the
batched GPU path's per-level row counts are checked.

level_rows[i] must equal the number of injective maps of the query vertices
vorder[0..i] into their candidate sets that preserve every query edge among
them (partial embeddings of the prefix), found here by an independent
backtracking search; the last level must equal orc_count and the brute-force
backtracker; the count cap must saturate exactly above the true count."""
import numpy as np
import pytest

import oracle
from gem_amd import Graph, Query


def _adjacency(g):
    return [set(g.neighbors[g.offsets[v]:g.offsets[v + 1]].tolist()) for v in range(g.num_vertices)]


def prefix_maps(adj, can, q, verts):
    """Injective maps of `verts` (a connected prefix, in order) into their
    candidate sets preserving the query edges among them."""
    pos = {u: i for i, u in enumerate(verts)}
    back = [[pos[b] for a, b in q.edges.tolist() if a == u and b in pos and pos[b] < i] +
            [pos[a] for a, b in q.edges.tolist() if b == u and a in pos and pos[a] < i] for i, u in enumerate(verts)]
    cand = [set(can[u].tolist()) for u in verts]
    img = [0] * len(verts)

    def rec(i):
        if i == len(verts):
            return 1
        if back[i]:
            pool = adj[img[back[i][0]]] & cand[i]
        else:
            pool = cand[i]
        used = set(img[:i])
        n = 0
        for v in pool:
            if v in used:
                continue
            if all(v in adj[img[j]] for j in back[i]):
                img[i] = v
                n += rec(i + 1)
        return n

    return rec(0)


def random_connected_order(q, rng):
    deg = np.bincount(q.edges.reshape(-1), minlength=q.num_v)
    start = int(rng.choice([u for u in range(q.num_v) if deg[u] != 1]))
    order, seen = [start], {start}
    nbrs = {u: set() for u in range(q.num_v)}
    for a, b in q.edges.tolist():
        nbrs[a].add(b)
        nbrs[b].add(a)
    while len(order) < q.num_v:
        front = sorted({w for u in order for w in nbrs[u]} - seen)
        w = int(rng.choice(front))
        order.append(w)
        seen.add(w)
    return np.array(order, np.uint32)


def _queries(g, n, sizes, seed):
    out = []
    i = 0
    while len(out) < n:
        q = Query.random_walk(g, sizes[i % len(sizes)], seed + i)
        i += 1
        if q.num_e >= 2:
            out.append(q)
    return out


@pytest.mark.parametrize("gseed", [3, 17])
def test_level_rows_are_prefix_embedding_counts(gseed):
    # two labels on a small dense graph: equal-label query vertices are
    # common, so a walk that forgot injectivity would overcount
    g = Graph.erdos_renyi(36, 150, 2, gseed)
    adj = _adjacency(g)
    rng = np.random.default_rng(gseed)
    checked = 0
    for q in _queries(g, 5, [4, 5, 6], 1000 * gseed):
        can = oracle.filter(g, q)
        for _ in range(3):
            vo = random_connected_order(q, rng)
            r = oracle.vcount(g, q, vo)
            for i in range(q.num_v):
                want = prefix_maps(adj, can, q, vo[:i + 1].tolist())
                assert int(r["level_rows"][i]) == want, (q.edges.tolist(), vo.tolist(), i)
            checked += 1
    assert checked == 15


@pytest.mark.parametrize("gseed", [5, 9, 23])
def test_final_level_equals_count_and_bruteforce(gseed):
    g = Graph.erdos_renyi(60, 260, 3, gseed)
    rng = np.random.default_rng(gseed)
    for q in _queries(g, 6, [4, 5, 6, 7], 50 * gseed):
        want = oracle.bruteforce(g, q)
        assert oracle.count(g, q)["count"] == want
        for _ in range(2):
            vo = random_connected_order(q, rng)
            r = oracle.vcount(g, q, vo)
            assert r["count"] == want and not r["saturated"]
            assert int(r["level_rows"][-1]) == want


def test_cap_saturates_exactly_above_the_count():
    g = Graph.erdos_renyi(80, 500, 2, 41)
    q = next(q for q in _queries(g, 20, [5], 7) if oracle.bruteforce(g, q) > 10)
    n = oracle.bruteforce(g, q)
    vo = random_connected_order(q, np.random.default_rng(1))
    at = oracle.vcount(g, q, vo, cap=n)
    assert at["count"] == n and not at["saturated"]
    below = oracle.vcount(g, q, vo, cap=n - 1)
    assert below["saturated"] and below["count"] == n - 1


def test_prefix_search_detects_a_dropped_injectivity_check():
    """The independent search is sensitive to what the pin is about: without
    the distinctness test it counts more on a same-label path query."""
    g = Graph.erdos_renyi(30, 120, 1, 2)
    q = Query(np.zeros(3, np.uint32), np.array([(0, 1), (1, 2)], np.uint32))
    can = oracle.filter(g, q)
    adj = _adjacency(g)
    exact = prefix_maps(adj, can, q, [1, 0, 2])
    non_injective = sum(len(adj[v]) ** 2 for v in can[1].tolist())
    assert exact == oracle.vcount(g, q, np.array([1, 0, 2], np.uint32))["count"]
    assert non_injective > exact
