"""Seeded synthetic workloads (data graph + query set) for the benchmark and
the parity tests.

  er10k  Erdos-Renyi, 10,000 vertices, 50,000 edges, 8 uniform labels;
         20 random-walk 5-vertex queries.
  rmat20 R-MAT 2^20 vertices, average degree 16 (a,b,c,d = .57,.19,.19,.05),
         16 Zipf(s=1) labels; 100 random-walk 8-vertex queries, 50 dense
         (average degree >= 3) and 50 sparse, by rejection sampling.
  rmat22 R-MAT 2^22 vertices, average degree 16, 32 Zipf labels; 512
         random-walk queries with sizes uniform in [4, 32], mixed density.
  rmat20_learn  the rmat20 graph with 256 queries of 8-24 vertices (the
         learned-order workload; held out from its model's training).
  rmat20_train  the rmat20 graph with other queries of 8-24 vertices
         (training set of the order model used on rmat20 / rmat20_learn).
  rmat22_learn  the rmat22 graph with other queries of 4-32 vertices
         (training set of the order model used on rmat22).
  ooc24  R-MAT 2^24 vertices, average degree 64, 64 Zipf labels; 32 sparse
         (average degree < 3) random-walk 12-vertex queries: the out-of-core
         shape (streamed edge tables) at a size one box generates in about a
         minute (2^28 vertices take over ten).
  ooc26  the same at 2^26 vertices (4.3e9 directed slots, 17 GB of
         neighbours): the largest out-of-core shape one command generates and
         runs within its time limit.
"""
import numpy as np

from .graph import Graph, Query

WORKLOADS = {
    "er10k": dict(kind="er", n=10000, m=50000, labels=8, seed=1001),
    "rmat20": dict(kind="rmat", scale=20, degree=16.0, labels=16, seed=1002),
    "rmat22": dict(kind="rmat", scale=22, degree=16.0, labels=32, seed=1003),
    "rmat20_learn": dict(kind="rmat", scale=20, degree=16.0, labels=16, seed=1002),
    "rmat20_train": dict(kind="rmat", scale=20, degree=16.0, labels=16, seed=1002, qseed=2002),
    "rmat22_learn": dict(kind="rmat", scale=22, degree=16.0, labels=32, seed=1003, qseed=2003),
    "ooc24": dict(kind="rmat", scale=24, degree=64.0, labels=64, seed=1005),
    "ooc26": dict(kind="rmat", scale=26, degree=64.0, labels=64, seed=1006),
}

CAP = 10 ** 9  # enumeration cap for the R-MAT workloads


def make_graph(name):
    w = WORKLOADS[name]
    if w["kind"] == "er":
        return Graph.erdos_renyi(w["n"], w["m"], w["labels"], w["seed"])
    return Graph.rmat(w["scale"], w["degree"], w["labels"], w["seed"])


def make_queries(name, graph, count=None):
    seed = WORKLOADS[name].get("qseed", WORKLOADS[name]["seed"])
    qs = []
    if name == "er10k":
        n = count or 20
        i = 0
        while len(qs) < n:
            q = Query.random_walk(graph, 5, seed * 1000 + i)
            i += 1
            if q.num_e >= 2:
                qs.append(q)
        return qs
    if name == "rmat20":
        n = count or 100
        dense, sparse = [], []
        i = 0
        while len(dense) < n // 2 or len(sparse) < n - n // 2:
            q = Query.random_walk(graph, 8, seed * 1000 + i)
            i += 1
            if q.dense and len(dense) < n // 2:
                dense.append(q)
            elif not q.dense and len(sparse) < n - n // 2:
                sparse.append(q)
        # interleave so any prefix is mixed
        for d, s in zip(dense, sparse):
            qs += [d, s]
        return qs
    if name in ("ooc24", "ooc26"):
        n = count or 32
        i = 0
        while len(qs) < n:
            q = Query.random_walk(graph, 12, seed * 1000 + i)
            i += 1
            if q.num_e >= 2 and not q.dense:
                qs.append(q)
        return qs
    lo, hi, n = {"rmat22": (4, 32, count or 512), "rmat22_learn": (4, 32, count or 256)}.get(
        name, (8, 24, count or 256))
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=n)
    i = 0
    for k in sizes:
        while True:
            q = Query.random_walk(graph, int(k), seed * 100000 + i)
            i += 1
            if q.num_e >= 2:
                qs.append(q)
                break
    return qs
