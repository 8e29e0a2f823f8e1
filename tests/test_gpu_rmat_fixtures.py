"""GPU counts on the R-MAT workloads against the oracle's fixtures.

tests/golden/make_rmat_fixtures.py wrote, with the CPU oracle, the count and
saturated flag (cap 1e9) of every query of the R-MAT 2^20 / 100-query
workload, every 16th query of the R-MAT 2^22 / 512-query batch and every 8th
query of the 8-24-vertex learned-order workload on the 2^20 graph.  Here the
same seeded generators rebuild the graphs and queries (checked against the
recorded digests), and the batched path (model-ranked orders with restarts
and racing twins, as the benchmark runs it; without the twins; without the
model), and the plain
path on a sample must reproduce every recorded result.  Queries the oracle
did not finish are listed in the fixture and skipped here.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from gem_amd.learned import OrderModel, match_batch_learned
from gem_amd.workloads import make_graph, make_queries

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIRST_ATTEMPT = 2 * 10 ** 9


def _fixture(name):
    path = os.path.join(HERE, "golden", f"{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"no fixture {name}.json")
    return json.load(open(path))


def _digest(g):
    h = hashlib.sha256()
    for a in (g.offsets, g.neighbors, g.labels):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _setup(ctx, name):
    fx = _fixture(name)
    g = make_graph(name)
    assert _digest(g) == fx["graph"]["sha256"], "graph generator changed since the fixture was written"
    qs = make_queries(name, g)
    assert len(qs) == fx["num_queries"]
    picked = []
    for e in fx["queries"]:
        q = qs[e["index"]]
        assert hashlib.sha1(q.labels.tobytes() + q.edges.tobytes()).hexdigest() == e["labels_sha1"]
        if not e.get("unfinished"):
            picked.append((e, q))
    ctx.load_data(g)
    return fx, g, picked


def _model(ctx, graph_name):
    path = os.path.join(ROOT, "gem_amd", "models", f"order_mlp_{graph_name}.npz")
    return OrderModel.load(ctx, path) if os.path.exists(path) else None


@pytest.mark.parametrize("name,model_name", [("rmat20", "rmat20"), ("rmat22", "rmat22"),
                                             ("rmat20_learn", "rmat20")])
def test_batched_counts_match_oracle_fixtures(gpu_ctx, name, model_name):
    fx, g, picked = _setup(gpu_ctx, name)
    print(f"{name}: {len(picked)} fixture queries", flush=True)
    qs = [q for _, q in picked]
    model = _model(gpu_ctx, model_name)
    for m in (model, None):
        res = match_batch_learned(gpu_ctx, qs, m, max_orders=8, cap=fx["cap"], first_attempt=FIRST_ATTEMPT)
        bad = [(e["index"], (e["count"], e["saturated"]), (r["count"], r["saturated"], r["work_limited"]))
               for (e, _), r in zip(picked, res)
               if r["work_limited"] or (r["count"], r["saturated"]) != (e["count"], e["saturated"])]
        assert not bad, bad
        if m is not None:
            m.close()


def test_racing_off_matches_oracle_fixtures(gpu_ctx):
    """Without twins racing the stragglers (and with them, above) the R-MAT
    2^20 workload's counts equal the fixture."""
    fx, g, picked = _setup(gpu_ctx, "rmat20")
    qs = [q for _, q in picked]
    model = _model(gpu_ctx, "rmat20")
    gpu_ctx.set_racing(False)
    try:
        res = match_batch_learned(gpu_ctx, qs, model, max_orders=8, cap=fx["cap"], first_attempt=10 ** 7)
    finally:
        gpu_ctx.set_racing(True)
        if model is not None:
            model.close()
    bad = [(e["index"], r["count"], r["saturated"]) for (e, _), r in zip(picked, res)
           if r["work_limited"] or (r["count"], r["saturated"]) != (e["count"], e["saturated"])]
    assert not bad, bad


@pytest.mark.parametrize("name", ["rmat20", "rmat22"])
def test_plain_counts_match_oracle_fixtures(gpu_ctx, name):
    """The plain per-query path (GEM edge order) on a sample of the fixture
    queries, under a work limit of 2e9 rows read; every query it finishes
    must agree."""
    fx, g, picked = _setup(gpu_ctx, name)
    gpu_ctx.set_work_limit(2 * 10 ** 9)
    try:
        finished = 0
        for e, q in picked[::max(1, len(picked) // 6)][:6]:
            print(f"{name}: plain query {e['index']}", flush=True)
            r = gpu_ctx.match_full(q, cap=fx["cap"])
            if r["work_limited"]:
                continue
            finished += 1
            assert (r["count"], r["saturated"]) == (e["count"], e["saturated"]), e["index"]
        assert finished >= 1
    finally:
        gpu_ctx.set_work_limit(0)


def test_streamed_counts_match_oracle_fixtures(gpu_ctx):
    """The R-MAT 2^20 workload's 100 queries with their edge tables streamed
    from pinned host memory through a 256 MiB device cache: every count equals
    the oracle's fixture."""
    fx, g, picked = _setup(gpu_ctx, "rmat20")
    qs = [q for _, q in picked]
    model = _model(gpu_ctx, "rmat20")
    try:
        gpu_ctx.set_table_stream(256 << 20, 1 << 20)
        res = match_batch_learned(gpu_ctx, qs, model, max_orders=8, cap=fx["cap"], first_attempt=FIRST_ATTEMPT)
        st = gpu_ctx.stream_stats()
    finally:
        gpu_ctx.set_table_stream(0)
        if model is not None:
            model.close()
    print("stream stats", st, flush=True)
    assert st["table_bytes"] > (256 << 20)
    bad = [(e["index"], r["count"], r["saturated"], r["work_limited"]) for (e, _), r in zip(picked, res)
           if r["work_limited"] or (r["count"], r["saturated"]) != (e["count"], e["saturated"])]
    assert not bad, bad


@pytest.mark.parametrize("overlap", [True, False])
def test_prep_overlap_matches_oracle_fixtures(gpu_ctx, overlap):
    """With 8 queries running at once, the R-MAT 2^20 workload's queries past
    the first 24 are prepared on the second stream while the join runs (or,
    with the overlap off, all before it): every count equals the fixture."""
    fx, g, picked = _setup(gpu_ctx, "rmat20")
    qs = [q for _, q in picked]
    assert len(qs) > 24
    model = _model(gpu_ctx, "rmat20")
    gpu_ctx.set_max_active(8)
    gpu_ctx.set_prep_overlap(overlap)
    try:
        res = match_batch_learned(gpu_ctx, qs, model, max_orders=8, cap=fx["cap"], first_attempt=FIRST_ATTEMPT)
    finally:
        gpu_ctx.set_prep_overlap(True)
        gpu_ctx.set_max_active(64)
        if model is not None:
            model.close()
    bad = [(e["index"], r["count"], r["saturated"], r["work_limited"]) for (e, _), r in zip(picked, res)
           if r["work_limited"] or (r["count"], r["saturated"]) != (e["count"], e["saturated"])]
    assert not bad, bad
