"""Writes the R-MAT count fixtures of this directory with the CPU oracle
(run from the repo root; takes tens of minutes on 8 cores):

    python tests/golden/make_rmat_fixtures.py [rmat20|rmat22|rmat20_learn ...]

rmat20.json        all 100 queries of the R-MAT 2^20 / 16-label workload
                   (8-vertex queries, half dense, half sparse)
rmat22.json        every 16th of the 512 queries of the R-MAT 2^22 / 32-label
                   batch (4-32 vertices), plus every 4-vertex query (the
                   ones whose counts stay below the cap)
rmat20_learn.json  every 8th of the 256 queries of 8-24 vertices on the
                   R-MAT 2^20 graph (the learned-order workload)

Per query: count and saturated flag (count capped at 1e9: a query with more
embeddings reports saturated and count = cap), from the oracle's vertex-level
depth-first count (orc_vcount) along an order picked here (most edges back to
the vertices already placed, then fewest candidates); the count does not depend
on the order.  A query the oracle does not finish within WORK_LIMIT
intermediate rows is listed with "unfinished": true and no count.  Each file
also records a digest of the generated graph, so a change of generator is
caught before counts are compared.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gem_amd.workloads import CAP, make_graph, make_queries  # noqa: E402

WORK_LIMIT = 4 * 10 ** 9
SAMPLE = {"rmat20": 1, "rmat22": 16, "rmat20_learn": 8}


def graph_digest(g):
    h = hashlib.sha256()
    for a in (g.offsets, g.neighbors, g.labels):
        h.update(np.ascontiguousarray(a).tobytes())
    return {"num_vertices": int(g.num_vertices), "num_edges": int(g.num_edges), "sha256": h.hexdigest()}


def pick_order(q, ncan):
    """Most edges back to the placed vertices first, then fewest candidates,
    then lowest id; starts at the vertex with the fewest candidates."""
    adj = [set() for _ in range(q.num_v)]
    for a, b in q.edges:
        adj[int(a)].add(int(b))
        adj[int(b)].add(int(a))
    order = [min(range(q.num_v), key=lambda u: (ncan[u], u))]
    placed = set(order)
    while len(order) < q.num_v:
        best = max((u for u in range(q.num_v) if u not in placed and adj[u] & placed),
                   key=lambda u: (len(adj[u] & placed), -ncan[u], -u))
        order.append(best)
        placed.add(best)
    return np.array(order, np.uint32)


def make(name):
    t0 = time.time()
    g = make_graph(name)
    qs = make_queries(name, g)
    step = SAMPLE[name]
    out = {"workload": name, "graph": graph_digest(g), "num_queries": len(qs), "sample_step": step, "cap": CAP,
           "work_limit": WORK_LIMIT, "queries": []}
    picked = sorted(set(range(0, len(qs), step)) | ({i for i, q in enumerate(qs) if q.num_v == 4}
                                                    if name == "rmat22" else set()))
    for i in picked:
        q = qs[i]
        tq = time.time()
        ncan = [len(c) for c in oracle.filter(g, q)]
        vo = pick_order(q, ncan)
        r = oracle.vcount(g, q, vo, cap=CAP, work_limit=WORK_LIMIT)
        e = {"index": i, "num_v": int(q.num_v), "num_e": int(q.num_e), "dense": bool(q.dense),
             "labels_sha1": hashlib.sha1(q.labels.tobytes() + q.edges.tobytes()).hexdigest()}
        if r is None:
            e["unfinished"] = True
        else:
            e.update(count=int(r["count"]), saturated=bool(r["saturated"]))
        e["oracle_s"] = round(time.time() - tq, 2)
        out["queries"].append(e)
        print(f"[{name}] query {i}: {e}", flush=True)
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"[{name}] wrote {len(out['queries'])} queries in {time.time() - t0:.0f}s "
          f"({sum(1 for e in out['queries'] if e.get('unfinished'))} unfinished)", flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or ["rmat20", "rmat20_learn", "rmat22"]:
        make(name)
