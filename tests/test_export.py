"""Canonical embedding files on the host: order, digest, header, and the
multi-run merge (tiny runs) against one in-memory sort."""
import numpy as np
import pytest

from gem_amd import GemError
from gem_amd.export import export_rows, read_embeddings


def _fnv1a(data, h=0xcbf29ce484222325):
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("nv", [1, 3, 8])
def test_rows_sorted_and_digested(tmp_path, nv):
    rng = np.random.default_rng(nv)
    rows = rng.integers(0, 50, size=(3000, nv), dtype=np.uint32)  # small range: ties on leading columns
    r1 = export_rows(rows, tmp_path / "one.emb")
    got, head = read_embeddings(tmp_path / "one.emb")
    want = rows[np.lexsort(rows.T[::-1])]
    assert np.array_equal(got, want)
    assert head == {"rows": 3000, "num_v": nv, "digest": r1["digest"]}
    assert r1["runs"] == 1 and r1["digest"] == _fnv1a(want.astype("<u4").tobytes())
    # runs of 97 rows: 31 sorted runs merged from temporary files
    r2 = export_rows(rows, tmp_path / "many.emb", run_bytes=97 * 4 * nv)
    assert r2["runs"] == 31 and r2["digest"] == r1["digest"]
    assert (tmp_path / "many.emb").read_bytes() == (tmp_path / "one.emb").read_bytes()
    assert sorted(p.name for p in tmp_path.iterdir()) == ["many.emb", "one.emb"]  # runs removed


def test_empty_and_invalid(tmp_path):
    r = export_rows(np.zeros((0, 4), np.uint32), tmp_path / "e.emb")
    rows, head = read_embeddings(tmp_path / "e.emb")
    assert rows.shape == (0, 4) and head["digest"] == 0xcbf29ce484222325 and r["rows"] == 0
    with pytest.raises(GemError):
        export_rows(np.zeros((2, 40), np.uint32), tmp_path / "wide.emb")  # more columns than GEM_MAX_QV
