"""The filter's no-allocation entry point (gem_hip_filter_noalloc) against the
allocating one (gem_hip_filter): candidate counts and lists bit-equal, on two
calls in a row (the second with a smaller query, so stale scratch contents
from the first would show), with the library's device pool unchanged by the
calls, and a clear error for a graph whose scratch need exceeds the fixed
array.  The test allocates the device output arrays itself with hipMalloc
from the HIP runtime the library is linked against (plumbing only)."""
import ctypes as C

import numpy as np
import pytest

from gem_amd import Graph, Query
from gem_amd._lib import GemError, load

pytestmark = pytest.mark.gpu

GEM_E_LIMIT = 7


def _hub_graph():
    """20,000 vertices, 3 labels, ~40,000 random edges plus 4 hubs of degree
    5,000-6,000 (above the filter's 4,096-entry hub chunk)."""
    rng = np.random.default_rng(91)
    n = 20000
    pairs = set()
    while len(pairs) < 40000:
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    for h in range(4):
        for b in rng.choice(np.arange(4, n), size=int(rng.integers(5000, 6000)), replace=False):
            pairs.add((h, int(b)))
    labels = rng.integers(0, 3, n).astype(np.uint32)
    return Graph.from_edges(n, sorted(pairs), labels)


def _hip():
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


def _dev_alloc(rt, nbytes):
    p = C.c_void_p()
    assert rt.hipMalloc(C.byref(p), nbytes) == 0
    assert rt.hipMemset(p, 0xFF, nbytes) == 0  # not zero: a list written short would show
    return p


def _noalloc(ctx, q, pitch):
    rt = _hip()
    lib = load()
    num = _dev_alloc(rt, 4 * q.num_v)
    can = _dev_alloc(rt, 4 * q.num_v * pitch)
    try:
        assert rt.hipDeviceSynchronize() == 0
        before = ctx.memory_stats()
        v = q.view()
        st = lib.gem_hip_filter_noalloc(ctx._h, C.byref(v), 0, num, can, pitch)
        ctx._check(st, "gem_hip_filter_noalloc")
        after = ctx.memory_stats()
        assert (after["in_use"], after["held"]) == (before["in_use"], before["held"])
        n = np.zeros(q.num_v, np.uint32)
        c = np.zeros(q.num_v * pitch, np.uint32)
        assert rt.hipMemcpy(n.ctypes.data_as(C.c_void_p), num, n.nbytes, 2) == 0  # device to host
        assert rt.hipMemcpy(c.ctypes.data_as(C.c_void_p), can, c.nbytes, 2) == 0
    finally:
        rt.hipFree(num)
        rt.hipFree(can)
    c = c.reshape(q.num_v, pitch)
    return n, [c[u, :n[u]].copy() for u in range(q.num_v)]


def test_noalloc_filter_equals_filter(gpu_ctx):
    g = _hub_graph()
    gpu_ctx.load_data(g)
    pitch = int(np.bincount(g.labels).max())
    big = Query.random_walk(g, 12, 3)
    small = Query.random_walk(g, 4, 5)
    hub = Query(np.array([g.labels[0], g.labels[5], g.labels[6]], np.uint32), np.array([[0, 1], [0, 2]], np.uint32))
    assert big.num_v > small.num_v
    for q in (big, small, hub):
        print("noalloc filter query", q.labels.tolist(), q.edges.tolist(), flush=True)
        want_n, want_l = gpu_ctx.filter(q)
        got_n, got_l = _noalloc(gpu_ctx, q, pitch)
        assert np.array_equal(got_n, want_n)
        for u in range(q.num_v):
            assert np.array_equal(got_l[u], want_l[u])


def test_noalloc_filter_rejects_oversize(gpu_ctx):
    # 2^23 isolated vertices and a 32-vertex path: 96 MiB of bitmaps, counts,
    # ranks and heavy list, beyond the 64 MiB array
    n = 1 << 23
    g = Graph.from_csr(np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(n, np.uint32))
    gpu_ctx.load_data(g)
    q = Query(np.zeros(32, np.uint32), np.array([[i, i + 1] for i in range(31)], np.uint32))
    with pytest.raises(GemError) as e:
        _noalloc(gpu_ctx, q, 1)
    assert e.value.status == GEM_E_LIMIT and "scratch" in str(e.value)
