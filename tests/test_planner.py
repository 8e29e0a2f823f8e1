"""GEM ordering: the product's closed-form planner (no lattice) against the
oracle's faithful restatement of the reference's lattice ids and scores."""
import numpy as np
import pytest

import oracle
from gem_amd import Graph, Query, gem_order


def _rand_connected_query(rng, nv, extra):
    edges = set()
    for v in range(1, nv):
        u = int(rng.integers(v))
        edges.add((u, v))
    tries = 0
    while len(edges) < nv - 1 + extra and tries < 1000:
        a, b = sorted(rng.choice(nv, 2, replace=False).tolist())
        edges.add((a, b))
        tries += 1
    e = sorted(edges)
    perm = rng.permutation(len(e))  # file order is arbitrary
    return Query(np.zeros(nv, np.uint32), np.array([e[i] for i in perm], np.uint32))


@pytest.mark.parametrize("seed", range(6))
def test_planner_matches_lattice_random_sizes(seed):
    rng = np.random.default_rng(seed)
    for _ in range(25):
        q = _rand_connected_query(rng, int(rng.integers(3, 8)), int(rng.integers(0, 5)))
        sizes = rng.integers(0, 1000, q.num_e).astype(np.uint64)
        want = oracle.gem_order(q, sizes)
        assert want is not None
        assert gem_order(q, sizes).tolist() == want.tolist(), (q.edges.tolist(), sizes.tolist())


@pytest.mark.parametrize("seed", range(6))
def test_planner_matches_lattice_with_ties(seed):
    """Equal table sizes make every choice a tie: the lattice-id tie-break
    decides the whole order."""
    rng = np.random.default_rng(100 + seed)
    for _ in range(25):
        q = _rand_connected_query(rng, int(rng.integers(3, 9)), int(rng.integers(0, 6)))
        sizes = rng.integers(0, 3, q.num_e).astype(np.uint64)
        want = oracle.gem_order(q, sizes)
        assert gem_order(q, sizes).tolist() == want.tolist(), (q.edges.tolist(), sizes.tolist())


def test_planner_matches_lattice_dense():
    rng = np.random.default_rng(7)
    for _ in range(5):
        q = _rand_connected_query(rng, 7, 9)  # 15 edges
        sizes = np.full(q.num_e, 5, np.uint64)
        want = oracle.gem_order(q, sizes)
        assert gem_order(q, sizes).tolist() == want.tolist()


def test_planner_on_random_walk_queries():
    g = Graph.erdos_renyi(500, 3000, 3, 3)
    for i in range(30):
        q = Query.random_walk(g, 4 + i % 5, 10 + i)
        if q.num_e < 2:
            continue
        ts = oracle.table_sizes(g, q)
        sizes = np.maximum(ts[0::2], ts[1::2])
        assert gem_order(q, sizes).tolist() == oracle.gem_order(q, sizes).tolist()


def test_first_edge_rule():
    # path 0-1-2-3: the middle edge has 2 adjacent edges, the ends 1.  With
    # equal sizes the connectivity term picks the middle edge (index 1).
    q = Query(np.zeros(4, np.uint32), np.array([(0, 1), (1, 2), (2, 3)], np.uint32))
    assert gem_order(q, np.array([10, 10, 10], np.uint64))[0] == 1
    # a much smaller end table outweighs it: 0.6/(2+0) > 0.6/(2+10) + 0.01
    assert gem_order(q, np.array([0, 10, 10], np.uint64))[0] == 0


def test_big_query_beyond_lattice():
    """Orders for queries past the reference's 32-edge bitset are defined and
    connected."""
    rng = np.random.default_rng(3)
    q = _rand_connected_query(rng, 24, 30)
    sizes = rng.integers(0, 4, q.num_e).astype(np.uint64)
    order = gem_order(q, sizes)
    assert sorted(order.tolist()) == list(range(q.num_e))
    seen = set(q.edges[order[0]].tolist())
    for e in order[1:]:
        assert set(q.edges[e].tolist()) & seen
        seen |= set(q.edges[e].tolist())
