"""ORACLE — test infrastructure only (the checker).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  It wraps oracle/liboracle.so, a C++/OpenMP restatement of the
reference's filter -> edge tables -> GEM order -> join path (see
oracle/gem_oracle.cpp for the file:line citations and pins).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_DIR, "liboracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", _DIR], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_PATH):
        build()
    lib = C.CDLL(_PATH)
    G = [C.c_uint32, u64p, u32p, u32p]
    Qv = [C.c_uint32, C.c_uint32, u32p, u32p]
    lib.orc_max_threads.restype = C.c_int
    lib.orc_filter.argtypes = G + Qv + [C.c_uint32, u32p, u32p, C.c_uint64, C.c_int]
    lib.orc_table_sizes.argtypes = G + Qv + [u64p, C.c_int]
    lib.orc_edge_table.argtypes = G + Qv + [C.c_uint32, u64p, u32p, u32p, C.c_int]
    lib.orc_gem_order.argtypes = [C.c_uint32, C.c_uint32, u32p, u64p, u32p, C.c_uint64]
    lib.orc_count.argtypes = G + Qv + [u32p, C.c_uint64, C.c_uint64, u64p, C.POINTER(C.c_int), u64p, u32p,
                                       C.POINTER(C.c_double), C.c_int]
    lib.orc_enumerate.argtypes = G + Qv + [u32p, u32p, C.c_uint64, u64p, C.c_int]
    lib.orc_vcount.argtypes = G + Qv + [u32p, C.c_uint64, C.c_uint64, u64p, C.POINTER(C.c_int), u64p, C.c_int]
    lib.orc_vcount.restype = C.c_int
    lib.orc_bruteforce.argtypes = G + Qv + [u64p]
    for f in ("orc_filter", "orc_table_sizes", "orc_edge_table", "orc_gem_order", "orc_count", "orc_enumerate",
              "orc_bruteforce"):
        getattr(lib, f).restype = C.c_int
    _lib = lib
    return lib


def _p32(a):
    return a.ctypes.data_as(u32p) if a is not None else None


def _p64(a):
    return a.ctypes.data_as(u64p) if a is not None else None


def _g(graph):
    off = np.ascontiguousarray(graph.offsets, dtype=np.uint64)
    nbr = np.ascontiguousarray(graph.neighbors, dtype=np.uint32)
    lab = np.ascontiguousarray(graph.labels, dtype=np.uint32)
    return (off, nbr, lab), [C.c_uint32(len(lab)), _p64(off), _p32(nbr), _p32(lab)]


def _q(query):
    lab = np.ascontiguousarray(query.labels, dtype=np.uint32)
    edg = np.ascontiguousarray(np.asarray(query.edges, dtype=np.uint32).reshape(-1))
    return (lab, edg), [C.c_uint32(len(lab)), C.c_uint32(len(edg) // 2), _p32(lab), _p32(edg)]


def max_threads():
    return int(_load().orc_max_threads())


def filter(graph, query, threshold=0, nthreads=0):
    """Can(u) lists (numpy arrays, ascending)."""
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    labs = np.asarray(graph.labels)
    pitch = max([int(np.count_nonzero(labs == l)) for l in query.labels] + [1])
    num = np.zeros(query.num_v, np.uint32)
    can = np.zeros((query.num_v, pitch), np.uint32)
    st = _load().orc_filter(*ga, *qa, threshold, _p32(num), _p32(can), pitch, nthreads)
    assert st == 0, st
    return [can[u, :num[u]].copy() for u in range(query.num_v)]


def table_sizes(graph, query, nthreads=0):
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    out = np.zeros(2 * query.num_e, np.uint64)
    st = _load().orc_table_sizes(*ga, *qa, _p64(out), nthreads)
    assert st == 0, st
    return out


def edge_table(graph, query, d, num_rows, size, nthreads=0):
    """(offsets, vertex ids, positions) of directed table d, or None if absent."""
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    off = np.zeros(num_rows + 1, np.uint64)
    ids = np.zeros(max(int(size), 1), np.uint32)
    pos = np.zeros(max(int(size), 1), np.uint32)
    st = _load().orc_edge_table(*ga, *qa, d, _p64(off), _p32(ids), _p32(pos), nthreads)
    if st == 1:
        return None
    assert st == 0, st
    return off, ids[:int(size)], pos[:int(size)]


def gem_order(query, edge_sizes, max_nodes=0):
    """GEM order through the reference's lattice (None if the lattice is too big)."""
    edg = np.ascontiguousarray(np.asarray(query.edges, dtype=np.uint32).reshape(-1))
    sz = np.ascontiguousarray(edge_sizes, dtype=np.uint64)
    out = np.zeros(query.num_e, np.uint32)
    st = _load().orc_gem_order(query.num_v, query.num_e, _p32(edg), _p64(sz), _p32(out), max_nodes)
    if st == 2:
        return None
    assert st == 0, st
    return out


def count(graph, query, order=None, cap=0, work_limit=0, nthreads=0):
    """dict(count, saturated, step_rows, order, ms=[filter, tables, join]) or
    None when work_limit intermediate rows were exceeded."""
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    ordv = None if order is None else np.ascontiguousarray(order, dtype=np.uint32)
    cnt = C.c_uint64()
    sat = C.c_int()
    steps = np.zeros(query.num_e, np.uint64)
    used = np.zeros(query.num_e, np.uint32)
    ms = (C.c_double * 3)()
    st = _load().orc_count(*ga, *qa, _p32(ordv), cap, work_limit, C.byref(cnt), C.byref(sat), _p64(steps),
                           _p32(used), ms, nthreads)
    if st == 2:
        return None
    assert st == 0, st
    return {"count": int(cnt.value), "saturated": bool(sat.value), "step_rows": steps, "order": used,
            "ms": [ms[0], ms[1], ms[2]]}


def vcount(graph, query, vorder, cap=0, work_limit=0, nthreads=0):
    """Vertex-level counts along vorder: dict(count, saturated, level_rows) or
    None at the work limit."""
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    vo = np.ascontiguousarray(vorder, dtype=np.uint32)
    cnt = C.c_uint64()
    sat = C.c_int()
    lr = np.zeros(query.num_v, np.uint64)
    st = _load().orc_vcount(*ga, *qa, _p32(vo), cap, work_limit, C.byref(cnt), C.byref(sat), _p64(lr), nthreads)
    if st == 2:
        return None
    assert st == 0, st
    return {"count": int(cnt.value), "saturated": bool(sat.value), "level_rows": lr}


def vertex_order_from_edges(query, order):
    """First-appearance vertex order of an edge order, starting at the first
    edge's non-leaf endpoint (restates the product's rule for tests)."""
    deg = np.bincount(np.asarray(query.edges).reshape(-1), minlength=query.num_v)
    s, d = (int(x) for x in query.edges[order[0]])
    a = s if deg[s] != 1 else d
    vo = [a, d if a == s else s]
    for e in order[1:]:
        for v in (int(x) for x in query.edges[e]):
            if v not in vo:
                vo.append(v)
    return np.array(vo, np.uint32)


def enumerate_embeddings(graph, query, order, capacity=1 << 22, nthreads=0):
    """Sorted embeddings (rows of vertex ids, column u = image of query vertex u)."""
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    ordv = np.ascontiguousarray(order, dtype=np.uint32)
    out = np.zeros(capacity * query.num_v, np.uint32)
    n = C.c_uint64()
    st = _load().orc_enumerate(*ga, *qa, _p32(ordv), _p32(out), capacity, C.byref(n), nthreads)
    assert st == 0, st
    return out[:n.value * query.num_v].reshape(-1, query.num_v)


def bruteforce(graph, query):
    keep_g, ga = _g(graph)
    keep_q, qa = _q(query)
    c = C.c_uint64()
    st = _load().orc_bruteforce(*ga, *qa, C.byref(c))
    assert st == 0, st
    return int(c.value)


def relax(query, threshold):
    """Relaxed queries as the reference's relax() reaches them
    (in-core/lib/plan/src/relax.cpp:84-142): breadth-first from the full edge
    tag, a tag at distance >= threshold or with at most num_v - 1 edges is not
    expanded, every present edge is flipped off in index order, and a new tag
    is kept only if checkConnectivity (relax.cpp:57-81: union-find over all
    num_v vertices, one component) holds.  Queries of at most one edge give the
    empty set (relax.cpp:88-89).  Returns the set of edge masks (int, bit i =
    edge i).  Pure Python: small queries only."""
    nv, ne = query.num_v, query.num_e
    edges = [(int(a), int(b)) for a, b in query.edges]

    def connected(tag):
        if nv <= 1:
            return True
        parent = list(range(nv))

        def find(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x

        comps = nv
        for i in range(ne):
            if (tag >> i) & 1:
                a, b = find(edges[i][0]), find(edges[i][1])
                if a != b:
                    parent[a] = b
                    comps -= 1
        return comps == 1

    out = set()
    if ne <= 1:
        return out
    full = (1 << ne) - 1
    out.add(full)
    queue = [(full, 0)]
    head = 0
    while head < len(queue):
        tag, dist = queue[head]
        head += 1
        if dist >= threshold or bin(tag).count("1") <= nv - 1:
            continue
        for i in range(ne):
            if not (tag >> i) & 1:
                continue
            nxt = tag & ~(1 << i)
            if nxt in out:
                continue
            out.add(nxt)
            if connected(nxt):
                queue.append((nxt, dist + 1))
            else:
                out.discard(nxt)
    return out
