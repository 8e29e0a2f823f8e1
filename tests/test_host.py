"""Host-side library: generators, reference text format, query extraction,
and the C ABI's exported symbols (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from gem_amd import GemError, Graph, Query, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    lib = load()
    hdr = open(os.path.join(ROOT, "include", "gem_hip.h")).read()
    names = set(re.findall(r"\b(gem_\w+)\s*\(", hdr))
    assert len(names) > 20
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "gem_amd", "_lib", "libgem_hip.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (gem_\w+)", out))
    missing = names - exported
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)


def _csr_ok(g):
    off, nbr = g.offsets, g.neighbors
    n = g.num_vertices
    for v in range(n):
        row = nbr[off[v]:off[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)
        assert v not in row
    # symmetric
    src = np.repeat(np.arange(n), np.diff(off).astype(np.int64))
    a = set(zip(src.tolist(), nbr.tolist()))
    assert all((d, s) in a for s, d in list(a)[:5000])


def test_er_generator_deterministic_and_exact():
    g1 = Graph.erdos_renyi(1000, 5000, 8, 42)
    g2 = Graph.erdos_renyi(1000, 5000, 8, 42)
    g3 = Graph.erdos_renyi(1000, 5000, 8, 43)
    assert g1.num_edges == 5000
    assert np.array_equal(g1.neighbors, g2.neighbors) and np.array_equal(g1.labels, g2.labels)
    assert not np.array_equal(g1.neighbors, g3.neighbors)
    assert g1.labels.max() < 8
    _csr_ok(g1)


def test_rmat_generator():
    g = Graph.rmat(12, 16.0, 16, 7)
    h = Graph.rmat(12, 16.0, 16, 7)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.neighbors, h.neighbors)
    assert g.num_vertices == 4096
    assert g.num_slots <= 4096 * 16
    _csr_ok(g)
    # Zipf(s=1): label 0 is the most frequent, frequency ~ 1/H_16 = 0.296
    freq = np.bincount(g.labels, minlength=16) / g.num_vertices
    assert abs(freq[0] - 0.296) < 0.03 and freq[0] > freq[15]
    # skew: R-MAT has hubs
    deg = g.degrees()
    assert deg.max() > 10 * deg.mean()


def test_text_roundtrip(tmp_path):
    g = Graph.erdos_renyi(300, 1200, 5, 3)
    p = tmp_path / "data.graph"
    g.save(p)
    h = Graph.load(p)
    assert np.array_equal(g.offsets, h.offsets)
    assert np.array_equal(g.neighbors, h.neighbors)
    assert np.array_equal(g.labels, h.labels)
    first = p.read_text().splitlines()[0]
    assert first == "t 300 1200"


def test_loader_rejects_bad_degree(tmp_path):
    p = tmp_path / "bad.graph"
    p.write_text("t 3 2\nv 0 0 1\nv 1 0 1\nv 2 0 1\ne 0 1\ne 1 2\n")
    with pytest.raises(GemError):
        Graph.load(p)


def test_query_random_walk_and_io(tmp_path):
    g = Graph.erdos_renyi(2000, 10000, 4, 5)
    q = Query.random_walk(g, 8, 1)
    assert q.num_v == 8
    assert np.all(q.edges[:, 0] < q.edges[:, 1])
    # induced subgraph of distinct data vertices -> connected
    seen = {0}
    changed = True
    while changed:
        changed = False
        for s, d in q.edges:
            if (s in seen) != (d in seen):
                seen |= {int(s), int(d)}
                changed = True
    assert len(seen) == 8
    p = tmp_path / "Q_0.in"
    q.save(p)
    r = Query.load(p)
    assert np.array_equal(q.labels, r.labels) and np.array_equal(q.edges, r.edges)
    assert Query.random_walk(g, 8, 1).edges.tolist() == q.edges.tolist()


def test_library_has_no_cpu_fallback():
    """The product's matcher entry points go straight to HIP: without a GPU,
    creating a context fails instead of computing on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from gem_amd import GemContext
    with pytest.raises(GemError):
        GemContext(0)


def test_committed_tuning_is_within_the_api_limits():
    """The autotuner's committed choice is accepted by the setters it feeds
    (ranges of gem_hip_set_persistent_blocks / gem_hip_set_heavy_split) and
    every trial it kept reported the same counts as the first."""
    import json
    path = os.path.join(ROOT, "gem_amd", "tuning", "rmat22.json")
    d = json.load(open(path))
    best = d["best"]
    assert 1 <= best["persistent_blocks"] <= 1 << 20
    assert 32 <= best["heavy_t"] <= 1 << 24 and 64 <= best["heavy_chunk"] <= 1 << 20
    assert best["level_budget"] == 0 and best["max_orders"] >= 1
    assert all(t["counts_equal"] for t in d["trials"])
    assert min(t["seconds"] for t in d["trials"]) == d["best_seconds"]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run(["python3", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_order_ranking_and_restart_schedule():
    """The ranking rule the batched path applies (GEM's order, candidate 0,
    stays first unless a candidate is predicted cheaper by more than the
    margin), the restart schedule's work (attempt a runs the (a mod n)-th
    ranked order with a limit of first * growth^(a // width)), and the
    relative training labels, on hand-worked cases."""
    from gem_amd.learned import rank_orders, relative_targets, schedule_work

    Q = np.array([0, 0, 0, 1, 1])
    C = np.array([0, 1, 2, 0, 1])
    pred = np.array([10.0, 7.5, 11.0, 6.0, 5.8])
    # query 0: 7.5 < 10 - 1 -> candidate 1 first; query 1: 5.8 is not < 6 - 1 -> GEM's first
    assert rank_orders(pred, Q, C, margin=1.0) == {0: [1, 0, 2], 1: [0, 1]}
    assert rank_orders(pred, Q, C, margin=0.0) == {0: [1, 0, 2], 1: [1, 0]}
    # costs: order 0 needs 50, order 1 needs 5, order 2 never ends
    costs = {0: 50.0, 1: 5.0, 2: float("inf")}
    assert schedule_work(costs, [1, 0, 2], 10) == 5.0                 # first attempt ends
    assert schedule_work(costs, [2, 0, 1], 10, growth=4) == 10 + 40 + 5  # limits 10, 40, then order 1
    assert schedule_work(costs, [0, 2, 1], 10, growth=4, width=3) == 10 + 10 + 5
    assert schedule_work(costs, [2, 0], 10, growth=4) == 10 + 40 + 160 + 50  # 0 ends at limit 640
    # labels: log2(1 + work) relative to the query's cheapest order
    y = relative_targets(np.array([3.0, 15.0, 1.0, 7.0]), np.array([0, 0, 1, 1]))
    assert np.allclose(y, [0.0, 2.0, 0.0, 2.0])


def test_profiles_cover_the_walk_kernels():
    """bench.py's roofline line reads the per-launch HBM traffic and rocprofv3
    average of its dominant kernel from profiles/: both are recorded for every
    walk kernel of the rmat22 step."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for k in ("k_vext<write>", "k_vext<final>", "k_vext_heavy<write>", "k_vext_heavy<final>"):
        assert bench.profiled_traffic(k, "rmat22") > 0
        assert bench.profiled("avg_launch_us", k, "rmat22") > 0
        # the per-pass total is launches x average, as rocprofv3 summed it
        n, tot = bench.profiled("launches_per_pass", k, "rmat22"), bench.profiled("total_us_per_pass", k, "rmat22")
        assert abs(tot / n - bench.profiled("avg_launch_us", k, "rmat22")) < 0.1
    assert bench.profiled_traffic("k_vext<write>", "no-such-workload") is None


def test_benchmark_workload_is_the_described_batch():
    """The rmat22 workload is the 512-query batch the metric is quoted on:
    2^22 vertices, queries of 4-32 vertices, generated deterministically, and
    disjoint from the order model's training queries (rmat22_learn: same
    graph, other query seed)."""
    from gem_amd.workloads import make_graph, make_queries
    g = make_graph("rmat22")
    assert g.num_vertices == 1 << 22
    qs = make_queries("rmat22", g)
    assert len(qs) == 512 and min(q.num_v for q in qs) == 4 and max(q.num_v for q in qs) == 32
    again = make_queries("rmat22", g)
    assert all(np.array_equal(a.edges, b.edges) and np.array_equal(a.labels, b.labels) for a, b in zip(qs, again))
    learn = make_queries("rmat22_learn", g)
    key = {(tuple(q.labels.tolist()), tuple(map(tuple, q.edges.tolist()))) for q in qs}
    assert not any((tuple(q.labels.tolist()), tuple(map(tuple, q.edges.tolist()))) in key for q in learn)


def _fnv1a(data, h=0xcbf29ce484222325):
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_binary_csr_cache(tmp_path):
    """Round trip through the binary CSR cache, its digest (FNV-1a 64 of each
    1 MiB piece of offsets, neighbours and labels, then of the piece digests),
    and rejection of a flipped byte and of a truncated file."""
    g = Graph.erdos_renyi(3000, 9000, 5, 8)
    p = tmp_path / "g.bin"
    g.save_binary(p)
    h = Graph.load(p)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.neighbors, h.neighbors)
    assert np.array_equal(g.labels, h.labels)
    pieces = []
    for arr in (g.offsets.astype(np.uint64), g.neighbors.astype(np.uint32), g.labels.astype(np.uint32)):
        b = arr.tobytes()
        pieces += [b[i:i + (1 << 20)] for i in range(0, len(b), 1 << 20)]
    want = _fnv1a(b"".join(_fnv1a(x).to_bytes(8, "little") for x in pieces))
    assert g.digest() == want == h.digest()
    raw = bytearray(p.read_bytes())
    assert raw[:8] == b"GEMCSR01" and int.from_bytes(raw[32:40], "little") == want
    bad = bytearray(raw)
    bad[64 + 8 * 3001 + 17] ^= 0x40  # a neighbour byte
    (tmp_path / "bad.bin").write_bytes(bad)
    with pytest.raises(GemError):
        Graph.load(tmp_path / "bad.bin")
    (tmp_path / "short.bin").write_bytes(raw[:-4])
    with pytest.raises(GemError):
        Graph.load(tmp_path / "short.bin")


def _reference_loader_csr(n, edges):
    """The host CSR the reference's loader leaves for an edge list: u32 offsets
    from the degrees, and each edge line appended to both endpoints' lists in
    file order (in-core/lib/io/src/io.cu:169-199)."""
    deg = np.zeros(n, np.int64)
    for s, d in edges:
        deg[s] += 1
        deg[d] += 1
    off = np.zeros(n + 1, np.uint32)
    off[1:] = np.cumsum(deg)
    nbr = np.zeros(int(off[-1]), np.uint32)
    fill = np.zeros(n, np.int64)
    for s, d in edges:
        nbr[off[s] + fill[s]] = d
        fill[s] += 1
        nbr[off[d] + fill[d]] = s
        fill[d] += 1
    return off, nbr


def test_graph_from_reference_host_csr():
    """The reference loader's file-order CSR becomes the sorted CSR of the same
    graph; a repeated edge, a self-loop, an asymmetric list and a neighbour out
    of range are refused."""
    rng = np.random.default_rng(7)
    g = Graph.erdos_renyi(400, 1500, 6, 3)
    off, nbr = g.offsets, g.neighbors
    edges = [(v, int(w)) for v in range(400) for w in nbr[off[v]:off[v + 1]] if v < w]
    edges = [edges[i] for i in rng.permutation(len(edges))]
    edges = [(d, s) if rng.random() < 0.5 else (s, d) for s, d in edges]
    ro, rn = _reference_loader_csr(400, edges)
    assert any(np.any(np.diff(rn[ro[v]:ro[v + 1]].astype(np.int64)) < 0) for v in range(400))  # unsorted input
    h = Graph.from_csr32(ro, rn, g.labels)
    assert np.array_equal(h.offsets, g.offsets) and np.array_equal(h.neighbors, g.neighbors)
    assert np.array_equal(h.labels, g.labels) and h.digest() == g.digest()
    lab = np.zeros(4, np.uint32)
    for bad_edges in ([(0, 1), (1, 2), (0, 1)], [(0, 1), (2, 2)]):
        ro, rn = _reference_loader_csr(4, bad_edges)
        with pytest.raises(GemError):
            Graph.from_csr32(ro, rn, lab)
    ro, rn = _reference_loader_csr(4, [(0, 1), (1, 2)])
    rn_asym = rn.copy()
    rn_asym[np.where(rn_asym == 2)[0][0]] = 3  # 1 lists 3, 3 does not list 1
    with pytest.raises(GemError):
        Graph.from_csr32(ro, rn_asym, lab)
    rn_far = rn.copy()
    rn_far[0] = 9
    with pytest.raises(GemError):
        Graph.from_csr32(ro, rn_far, lab)
