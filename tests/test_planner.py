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


def test_next_edge_tie_break_by_lattice_id_star3():
    """Hand-derived from the reference's lattice (lattice.cpp:41-128, BFS from
    the full edge set, removing edges in index order, ids in discovery order)
    and find_best_edge_candidate (unifiedTrie.cu:939-1018, candidates in
    ascending lattice id, strict >).

    Star 0-1, 0-2, 0-3 (edges e0, e1, e2), equal table sizes.
      id0 {e0,e1,e2}; level 1: id1 {e1,e2}, id2 {e0,e2}, id3 {e0,e1};
      level 2: from id1: id4 {e2}, id5 {e1}; from id2: {e2} seen, id6 {e0}.
    First edge: every score ties (same size, same two parents) -> e0.
    Next: parents of {e0} (id6) are id2 {e0,e2} and id3 {e0,e1}; the tie goes
    to the lower id, id2, i.e. e2 -- not e1, which the lowest edge index would
    pick.  Then e1."""
    q = Query(np.zeros(4, np.uint32), np.array([(0, 1), (0, 2), (0, 3)], np.uint32))
    sizes = np.array([7, 7, 7], np.uint64)
    assert gem_order(q, sizes).tolist() == [0, 2, 1]
    assert oracle.gem_order(q, sizes).tolist() == [0, 2, 1]


def test_next_edge_tie_break_by_lattice_id_star4():
    """Star with four leaves (e0..e3 = 0-1, 0-2, 0-3, 0-4), equal sizes.
      level 1: id1 {e1,e2,e3}, id2 {e0,e2,e3}, id3 {e0,e1,e3}, id4 {e0,e1,e2};
      level 2: from id1: id5 {e2,e3}, id6 {e1,e3}, id7 {e1,e2};
               from id2: id8 {e0,e3}, id9 {e0,e2}; from id3: id10 {e0,e1};
      level 3: from id5: id11 {e3}, id12 {e2}; from id6: id13 {e1};
               from id8: id14 {e0}.
    First edge e0 (all tie).  Parents of {e0} (id14): id8 {e0,e3}, id9
    {e0,e2}, id10 {e0,e1} -> e3.  Parents of {e0,e3} (id8): id2 {e0,e2,e3},
    id3 {e0,e1,e3} -> e2.  Then e1: order [0, 3, 2, 1]."""
    q = Query(np.zeros(5, np.uint32), np.array([(0, 1), (0, 2), (0, 3), (0, 4)], np.uint32))
    sizes = np.full(4, 3, np.uint64)
    assert gem_order(q, sizes).tolist() == [0, 3, 2, 1]
    assert oracle.gem_order(q, sizes).tolist() == [0, 3, 2, 1]
    # a strictly smaller table still wins over the tie-break: e1 next
    sizes2 = np.array([3, 2, 3, 3], np.uint64)
    # first edge: e1 has the smaller table -> e1; parents of {e1} are
    # id6 {e1,e3}, id7 {e1,e2}, id10 {e0,e1} (equal sizes 3) -> e3, then from
    # {e1,e3} (id6): id1 {e1,e2,e3}, id3 {e0,e1,e3} -> e2, then e0
    assert gem_order(q, sizes2).tolist() == [1, 3, 2, 0]
    assert oracle.gem_order(q, sizes2).tolist() == [1, 3, 2, 0]


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
