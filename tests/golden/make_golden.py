"""Writes the golden fixtures of this directory (run from the repo root):

    python tests/golden/make_golden.py

* data.graph       Erdos-Renyi, 300 vertices, 1200 edges, 4 uniform labels
                   (seed 4242), in the reference's text format
* Q_<i>.in         12 random-walk queries of 3-6 vertices (seeds 5000+i)
* expected.json    per query: count (oracle), brute-force count (must agree),
                   GEM edge order (oracle lattice), candidate-list sizes,
                   per-directed-edge table sizes, per-step rows
* embeddings.npz   per query: the sorted embedding set (rows of vertex ids)

The counts are produced by the oracle and cross-checked against the
independent brute-force backtracker before anything is written.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gem_amd import Graph, Query  # noqa: E402


def main():
    g = Graph.erdos_renyi(300, 1200, 4, 4242)
    g.save(os.path.join(HERE, "data.graph"))
    expected, embs = [], {}
    i = 0
    while len(expected) < 12:
        q = Query.random_walk(g, 3 + len(expected) % 4, 5000 + i)
        i += 1
        if q.num_e < 2:
            continue
        k = len(expected)
        q.save(os.path.join(HERE, f"Q_{k}.in"))
        r = oracle.count(g, q)
        bf = oracle.bruteforce(g, q)
        assert r["count"] == bf, (k, r["count"], bf)
        can = oracle.filter(g, q)
        emb = oracle.enumerate_embeddings(g, q, r["order"])
        assert len(emb) == r["count"]
        embs[f"q{k}"] = emb
        expected.append({"query": f"Q_{k}.in", "count": r["count"], "bruteforce": bf,
                         "gem_order": [int(x) for x in r["order"]],
                         "num_can": [int(len(c)) for c in can],
                         "table_sizes": [int(x) for x in oracle.table_sizes(g, q)],
                         "step_rows": [int(x) for x in r["step_rows"]]})
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump({"graph": "data.graph", "queries": expected}, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "embeddings.npz"), **embs)
    print(f"wrote {len(expected)} queries")


if __name__ == "__main__":
    main()
