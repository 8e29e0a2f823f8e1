"""Embedding export on the GPU: the canonical file holds exactly the oracle's
sorted enumeration, its digest equals the host writer's for those rows, the
multi-run merge gives the same bytes, and max_rows stops the export with no
file left behind."""
import numpy as np
import pytest

import oracle
from gem_amd import GemError, Graph, Query
from gem_amd.export import export_rows, read_embeddings

pytestmark = pytest.mark.gpu


def test_export_equals_oracle_enumeration(gpu_ctx, tmp_path):
    g = Graph.erdos_renyi(2000, 9000, 4, 55)
    gpu_ctx.load_data(g)
    checked = 0
    for s in range(12):
        q = Query.random_walk(g, 4 + s % 3, 600 + s)
        want_n = oracle.count(g, q)["count"]
        if want_n > 300000:
            continue
        print("export query", q.num_v, q.edges.tolist(), want_n, flush=True)
        res = gpu_ctx.export_embeddings(q, tmp_path / "q.emb")
        rows, head = read_embeddings(tmp_path / "q.emb")
        want = oracle.enumerate_embeddings(g, q, oracle.count(g, q)["order"])
        assert np.array_equal(rows, want)
        assert res["rows"] == head["rows"] == want_n
        ref = export_rows(want, tmp_path / "ref.emb")
        assert res["digest"] == head["digest"] == ref["digest"]
        if want_n > 1000:
            multi = gpu_ctx.export_embeddings(q, tmp_path / "m.emb", run_bytes=4 * q.num_v * 997)
            assert multi["runs"] > 1 and multi["digest"] == res["digest"]
            assert (tmp_path / "m.emb").read_bytes() == (tmp_path / "q.emb").read_bytes()
            with pytest.raises(GemError):
                gpu_ctx.export_embeddings(q, tmp_path / "cut.emb", max_rows=want_n - 1)
            assert not (tmp_path / "cut.emb").exists()
            assert not any(p.name.startswith("cut.emb") for p in tmp_path.iterdir())
        checked += 1
    assert checked >= 6
