"""Host graphs and queries (numpy views over the C library's host CSR).

Mirrors the reference's loader layer (in-core/lib/io/src/io.cu:45-220,
in-core/lib/graph/include/cpuGraph.h:12-67): data graphs are undirected,
labelled, CSR with sorted adjacency; queries are small labelled graphs whose
edge i is the reference's edge-tag bit i.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import GEM_MAX_QE, GEM_MAX_QV, GEM_RQ_WORDS, GemError, QueryView, CsrView, load


def _check(st, what):
    if st != 0:
        raise GemError(st, what)


class Graph:
    """A host CSR graph owned by the C library."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        lib = load()
        self.num_vertices = int(lib.gem_graph_num_vertices(self._h))
        self.num_slots = int(lib.gem_graph_num_slots(self._h))
        n, m = self.num_vertices, self.num_slots
        self.offsets = np.ctypeslib.as_array(lib.gem_graph_offsets(self._h), shape=(n + 1,))
        self.neighbors = (np.ctypeslib.as_array(lib.gem_graph_neighbors(self._h), shape=(m,))
                          if m else np.zeros(0, np.uint32))
        self.labels = np.ctypeslib.as_array(lib.gem_graph_labels(self._h), shape=(n,))

    @property
    def num_edges(self):
        return self.num_slots // 2

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            load().gem_graph_free(h)
            self._h = None

    # ---- constructors
    @classmethod
    def erdos_renyi(cls, n, m, num_labels, seed):
        h = C.c_void_p()
        _check(load().gem_graph_generate_er(n, m, num_labels, seed, C.byref(h)), "gem_graph_generate_er")
        return cls(h)

    @classmethod
    def rmat(cls, scale, avg_degree, num_labels, seed, a=0.57, b=0.19, c=0.19, zipf=True):
        h = C.c_void_p()
        _check(load().gem_graph_generate_rmat(scale, avg_degree, a, b, c, num_labels, 1 if zipf else 0, seed,
                                              C.byref(h)), "gem_graph_generate_rmat")
        return cls(h)

    @classmethod
    def from_csr(cls, offsets, neighbors, labels):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        nbr = np.ascontiguousarray(neighbors, dtype=np.uint32)
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
        h = C.c_void_p()
        _check(load().gem_graph_from_csr(len(lab), off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         nbr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         lab.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h)),
               "gem_graph_from_csr")
        return cls(h)

    @classmethod
    def from_csr32(cls, offsets, neighbors, labels):
        """The reference loader's host CSR (u32 offsets, adjacency in edge-line
        order): offsets widened and every list sorted."""
        off = np.ascontiguousarray(offsets, dtype=np.uint32)
        nbr = np.ascontiguousarray(neighbors, dtype=np.uint32)
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
        h = C.c_void_p()
        _check(load().gem_graph_from_csr32(len(lab), off.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           nbr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           lab.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h)),
               "gem_graph_from_csr32")
        return cls(h)

    @classmethod
    def from_edges(cls, n, edges, labels):
        """Build from an undirected edge list (pairs); duplicates are an error."""
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        src = np.concatenate([e[:, 0], e[:, 1]])
        dst = np.concatenate([e[:, 1], e[:, 0]])
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        off = np.zeros(n + 1, np.uint64)
        np.add.at(off, src + 1, 1)
        off = np.cumsum(off).astype(np.uint64)
        return cls.from_csr(off, dst.astype(np.uint32), labels)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        _check(load().gem_graph_load(str(path).encode(), C.byref(h)), f"gem_graph_load({path})")
        return cls(h)

    def save(self, path):
        _check(load().gem_graph_save(self._h, str(path).encode()), f"gem_graph_save({path})")

    def save_binary(self, path):
        """Binary CSR cache; Graph.load recognises it."""
        _check(load().gem_graph_save_binary(self._h, str(path).encode()), f"gem_graph_save_binary({path})")

    def digest(self):
        """FNV-1a 64 digest of offsets, neighbours and labels."""
        return int(load().gem_graph_digest(self._h))

    def csr_view(self):
        return CsrView(self.num_vertices, self.num_slots,
                       self.offsets.ctypes.data_as(C.POINTER(C.c_uint64)),
                       self.neighbors.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.labels.ctypes.data_as(C.POINTER(C.c_uint32)))

    def degrees(self):
        return np.diff(self.offsets).astype(np.int64)


@dataclass
class Query:
    """num_v vertices with original labels; edges (src < dst) in edge-tag order."""
    labels: np.ndarray
    edges: np.ndarray  # shape (num_e, 2)

    def __post_init__(self):
        self.labels = np.ascontiguousarray(self.labels, dtype=np.uint32)
        self.edges = np.ascontiguousarray(np.asarray(self.edges, dtype=np.uint32).reshape(-1, 2))

    @property
    def num_v(self):
        return len(self.labels)

    @property
    def num_e(self):
        return len(self.edges)

    @property
    def avg_degree(self):
        return 2.0 * self.num_e / self.num_v

    @property
    def dense(self):
        return self.avg_degree >= 3.0

    def view(self):
        return QueryView(self.num_v, self.num_e, self.labels.ctypes.data_as(C.POINTER(C.c_uint32)),
                         self.edges.ctypes.data_as(C.POINTER(C.c_uint32)))

    @classmethod
    def random_walk(cls, graph, k, seed, max_e=GEM_MAX_QE):
        lab = np.zeros(k, np.uint32)
        edg = np.zeros(2 * max_e, np.uint32)
        ne = C.c_uint32()
        _check(load().gem_query_random_walk(graph._h, k, seed, lab.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            edg.ctypes.data_as(C.POINTER(C.c_uint32)), max_e, C.byref(ne)),
               "gem_query_random_walk")
        return cls(lab, edg[:2 * ne.value].reshape(-1, 2))

    @classmethod
    def load(cls, path):
        lab = np.zeros(GEM_MAX_QV, np.uint32)
        edg = np.zeros(2 * GEM_MAX_QE, np.uint32)
        nv, ne = C.c_uint32(), C.c_uint32()
        _check(load().gem_query_load(str(path).encode(), GEM_MAX_QV, GEM_MAX_QE, C.byref(nv),
                                     lab.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(ne),
                                     edg.ctypes.data_as(C.POINTER(C.c_uint32))), f"gem_query_load({path})")
        return cls(lab[:nv.value], edg[:2 * ne.value].reshape(-1, 2))

    def save(self, path):
        v = self.view()
        _check(load().gem_query_save(C.byref(v), str(path).encode()), f"gem_query_save({path})")

    def subquery(self, edge_mask):
        """The relaxed query keeping the edges whose bit is set in `edge_mask`
        (an int, bit i = edge i), same vertices and labels."""
        keep = [i for i in range(self.num_e) if (edge_mask >> i) & 1]
        return Query(self.labels.copy(), self.edges[keep])


def masks_to_ints(masks):
    """Rows of GEM_RQ_WORDS uint64 words -> Python ints (bit i = edge i)."""
    return [sum(int(w) << (64 * k) for k, w in enumerate(row)) for row in np.asarray(masks).reshape(-1, GEM_RQ_WORDS)]


def relaxed_queries(query, threshold):
    """Edge masks (ints, bit i = edge i) of the relaxed queries of `query` at
    `threshold`, in the library's order (gem_plan_relaxed_queries)."""
    v = query.view()
    n = C.c_uint32()
    _check(load().gem_plan_relaxed_queries(C.byref(v), threshold, 0, None, C.byref(n)), "gem_plan_relaxed_queries")
    masks = np.zeros((max(n.value, 1), GEM_RQ_WORDS), np.uint64)
    _check(load().gem_plan_relaxed_queries(C.byref(v), threshold, n.value,
                                           masks.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(n)),
           "gem_plan_relaxed_queries")
    return masks_to_ints(masks[:n.value])


def edge_bits(mask, num_e):
    """The reference's printed edge-tag string: bit i of the mask is character
    width-1-i, width max(32, num_e) (std::bitset<32> for queries of at most 32
    edges)."""
    width = max(32, num_e)
    return "".join("1" if (mask >> i) & 1 else "0" for i in range(width - 1, -1, -1))


def gem_order(query, edge_sizes):
    """GEM join order (edge indices) given |T_e| per undirected edge."""
    sz = np.ascontiguousarray(edge_sizes, dtype=np.uint64)
    out = np.zeros(query.num_e, np.uint32)
    v = query.view()
    _check(load().gem_plan_gem_order(C.byref(v), sz.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     out.ctypes.data_as(C.POINTER(C.c_uint32))), "gem_plan_gem_order")
    return out
