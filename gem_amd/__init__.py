"""gem_amd — MI355X (gfx950) subgraph matching with GEM's edge-centric pipeline.

The compute path is libgem_hip.so (hand-written HIP kernels behind a C ABI,
include/gem_hip.h); this package is its Python mirror.
"""
from ._lib import GEM_MAX_QE, GEM_MAX_QV, GemError, load
from .graph import Graph, Query, edge_bits, gem_order, relaxed_queries
from .matcher import NOT_FOUND, GemContext

__all__ = ["GEM_MAX_QE", "GEM_MAX_QV", "GemError", "load", "Graph", "Query", "gem_order", "relaxed_queries",
           "edge_bits", "GemContext", "NOT_FOUND"]
