"""Canonical embedding files (gem_export_rows / gem_hip_export_embeddings).

A file holds rows of num_v vertex ids (column u = image of query vertex u)
sorted ascending lexicographically: a 32-byte header ("GEMEMB01", u32 num_v,
u32 0, u64 rows, u64 FNV-1a 64 digest of the row bytes), then the rows as
little-endian u32.
"""
import ctypes as C

import numpy as np

from ._lib import ExportResult, GemError, load

MAGIC = b"GEMEMB01"


def _result(r):
    return {"rows": int(r.rows), "digest": int(r.digest), "runs": int(r.runs), "num_v": int(r.num_v)}


def export_rows(rows, path, run_bytes=0):
    """Write host rows (any order) as a canonical file."""
    a = np.ascontiguousarray(rows, dtype=np.uint32)
    if a.ndim != 2:
        raise ValueError("rows must be 2-D (n, num_v)")
    r = ExportResult()
    st = load().gem_export_rows(a.ctypes.data_as(C.POINTER(C.c_uint32)), a.shape[0], a.shape[1],
                                str(path).encode(), int(run_bytes), C.byref(r))
    if st:
        raise GemError(st, f"gem_export_rows({path})")
    return _result(r)


def read_embeddings(path):
    """(rows array (n, num_v) uint32, header dict)."""
    with open(path, "rb") as f:
        head = f.read(32)
        if len(head) != 32 or head[:8] != MAGIC:
            raise ValueError(f"{path}: not an embedding file")
        nv = int.from_bytes(head[8:12], "little")
        n = int.from_bytes(head[16:24], "little")
        digest = int.from_bytes(head[24:32], "little")
        rows = np.fromfile(f, dtype="<u4", count=n * nv).reshape(n, nv)
    if rows.shape[0] != n:
        raise ValueError(f"{path}: truncated")
    return rows, {"rows": n, "num_v": nv, "digest": digest}
