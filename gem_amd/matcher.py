"""GPU matcher: the reference's filter -> index -> tables -> join sequence
(in-core/src/main.cu:148-236) driven through the C ABI of libgem_hip.so.

Every call runs HIP kernels; there is no CPU path.
"""
import ctypes as C
import weakref

import numpy as np

from ._lib import (GEM_MAX_QV, GEM_RQ_WORDS, BatchOpts, GemError, JoinResult, KernelStat, MemoryStats, PhaseTimes, QueryView, StepLog,
                   StreamStats, load)

NOT_FOUND = 0xFFFFFFFF


def _p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class GemContext:
    """One device context (one per host thread), holding the uploaded data
    graph and the current query's candidates and edge tables."""

    def __init__(self, device=0):
        self._lib = load()
        h = C.c_void_p()
        st = self._lib.gem_hip_create(device, C.byref(h))
        if st != 0:
            raise GemError(st, f"gem_hip_create(device={device})")
        self._h = h
        self._models = weakref.WeakSet()
        self.graph = None
        self.query = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            for m in list(getattr(self, "_models", ())):
                m.close()
            self._lib.gem_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st, what):
        if st != 0:
            msg = self._lib.gem_hip_last_error(self._h)
            raise GemError(st, f"{what}: {msg.decode() if msg else ''}")

    def set_level_budget(self, nbytes):
        self._check(self._lib.gem_hip_set_level_budget(self._h, int(nbytes)), "gem_hip_set_level_budget")

    # ---- reference phases
    def load_data(self, graph):
        v = graph.csr_view()
        self._check(self._lib.gem_hip_load_data_csr(self._h, C.byref(v)), "gem_hip_load_data_csr")
        self.graph = graph

    def filter(self, query, threshold=0, want_lists=True):
        """filterSSM: returns |Can(u)| and (optionally) the sorted lists."""
        self.query = query
        num = np.zeros(query.num_v, np.uint32)
        v = query.view()
        if not want_lists:
            self._check(self._lib.gem_hip_filter(self._h, C.byref(v), threshold, _p32(num), None, 0),
                        "gem_hip_filter")
            return num, None
        # upper bound on list length: the data graph's label frequency
        freq = {}
        labs, cnts = np.unique(self.graph.labels, return_counts=True)
        for lab, cnt in zip(labs, cnts):
            freq[int(lab)] = int(cnt)
        pitch = max([freq.get(int(l), 0) for l in query.labels] + [1])
        buf = np.zeros((query.num_v, pitch), np.uint32)
        self._check(self._lib.gem_hip_filter(self._h, C.byref(v), threshold, _p32(num), _p32(buf), pitch),
                    "gem_hip_filter")
        return num, [buf[u, :num[u]].copy() for u in range(query.num_v)]

    def position_lookup(self, u, vertices):
        vs = np.ascontiguousarray(vertices, dtype=np.uint32)
        out = np.zeros(len(vs), np.uint32)
        self._check(self._lib.gem_hip_position_lookup(self._h, u, _p32(vs), len(vs), _p32(out)),
                    "gem_hip_position_lookup")
        return out

    def build_edge_tables(self):
        sizes = np.zeros(2 * self.query.num_e, np.uint64)
        self._check(self._lib.gem_hip_build_edge_tables(self._h, _p64(sizes)), "gem_hip_build_edge_tables")
        return sizes

    def read_edge_table(self, d, num_rows, size):
        off = np.zeros(num_rows + 1, np.uint64)
        val = np.zeros(max(int(size), 1), np.uint32)
        self._check(self._lib.gem_hip_read_edge_table(self._h, d, _p64(off), _p32(val)), "gem_hip_read_edge_table")
        return off, val[:int(size)]

    def join(self, order, cap=0, log=False, embeddings_capacity=0):
        order = np.ascontiguousarray(order, dtype=np.uint32)
        res = JoinResult()
        lg = StepLog() if log else None
        emb = (np.zeros(int(embeddings_capacity) * self.query.num_v, np.uint32)
               if embeddings_capacity else None)
        self._check(self._lib.gem_hip_join(self._h, _p32(order), cap, C.byref(res),
                                           C.byref(lg) if lg is not None else None,
                                           _p32(emb) if emb is not None else None,
                                           int(embeddings_capacity)), "gem_hip_join")
        out = {"count": int(res.count), "saturated": bool(res.saturated), "num_steps": int(res.num_steps)}
        if lg is not None:
            n = self.query.num_e
            out["step_rows"] = np.array(lg.rows[:n], dtype=np.uint64)
            out["step_ms"] = np.array(lg.ms[:n], dtype=np.float32)
            out["step_select"] = np.array(lg.is_select[:n], dtype=bool)
            out["step_edge"] = np.array(lg.edge[:n], dtype=np.uint32)
        if emb is not None:
            w = int(res.num_embeddings_written)
            out["embeddings"] = emb[:w * self.query.num_v].reshape(w, self.query.num_v) if w else None
        return out

    def set_profiling(self, on=True):
        self._check(self._lib.gem_hip_set_profiling(self._h, 1 if on else 0), "gem_hip_set_profiling")

    def memory_stats(self, reset_peak=False):
        """Device allocator figures (in_use, peak_in_use, held, reusable bytes)
        and the last batched run's level budget, peak charged level bytes,
        scheduler iterations and iterations forced past the budget."""
        st = MemoryStats()
        self._check(self._lib.gem_hip_memory_stats(self._h, 1 if reset_peak else 0, C.byref(st)),
                    "gem_hip_memory_stats")
        return {k: int(getattr(st, k)) for k, _ in MemoryStats._fields_}

    def kernel_stats(self):
        """{kernel name: dict(launches, ms, bytes)} accumulated while profiling."""
        arr = (KernelStat * 32)()
        n = C.c_uint32()
        self._check(self._lib.gem_hip_kernel_stats(self._h, arr, 32, C.byref(n)), "gem_hip_kernel_stats")
        return {arr[i].name.decode(): {"launches": int(arr[i].launches), "ms": float(arr[i].ms),
                                       "bytes": float(arr[i].bytes)} for i in range(n.value)}

    def phase_times(self):
        pt = PhaseTimes()
        self._check(self._lib.gem_hip_phase_times(self._h, C.byref(pt)), "gem_hip_phase_times")
        return {"filter_ms": pt.filter_ms, "index_ms": pt.index_ms, "tables_ms": pt.tables_ms,
                "join_ms": pt.join_ms}

    def set_persistent_blocks(self, blocks):
        self._check(self._lib.gem_hip_set_persistent_blocks(self._h, int(blocks)), "gem_hip_set_persistent_blocks")

    def set_heavy_split(self, threshold, chunk):
        """Rows whose driving list is longer than `threshold` are walked in
        `chunk`-entry pieces spread over waves (load balance only)."""
        self._check(self._lib.gem_hip_set_heavy_split(self._h, int(threshold), int(chunk)), "gem_hip_set_heavy_split")

    def set_level_swap(self, host_bytes):
        """Pending join levels may move to pinned host memory, up to host_bytes
        (least recently used first), when the device-resident ones exceed the
        level budget (0 = never)."""
        self._check(self._lib.gem_hip_set_level_swap(self._h, int(host_bytes)), "gem_hip_set_level_swap")

    def swap_stats(self):
        """Swapping statistics of the last batched run."""
        from ._lib import SwapStats
        st = SwapStats()
        self._check(self._lib.gem_hip_swap_stats(self._h, C.byref(st)), "gem_hip_swap_stats")
        return {k: int(getattr(st, k)) for k, _ in SwapStats._fields_}

    def set_racing(self, on=True):
        """Twins race the batched learned path's stragglers (default on)."""
        self._check(self._lib.gem_hip_set_racing(self._h, int(bool(on))), "gem_hip_set_racing")

    def set_prep_overlap(self, on=True):
        """The batched learned path prepares the queries past its first
        max_active + 16 on a second stream while the join runs (default on)."""
        self._check(self._lib.gem_hip_set_prep_overlap(self._h, int(bool(on))), "gem_hip_set_prep_overlap")

    def set_relaxed_sharing(self, on=True):
        """Relaxed queries share one filter at the threshold and one table per
        query edge (default), or are filtered and tabled one by one."""
        self._check(self._lib.gem_hip_set_relaxed_sharing(self._h, int(bool(on))), "gem_hip_set_relaxed_sharing")

    def set_graph_residency(self, host):
        """host=True: the next load_data keeps the neighbour array in pinned
        host memory (read over PCIe by the filter and table kernels)."""
        self._check(self._lib.gem_hip_set_graph_residency(self._h, int(bool(host))), "gem_hip_set_graph_residency")

    def set_max_active(self, queries):
        """Queries of a batched run holding join levels at once (0 = all)."""
        self._check(self._lib.gem_hip_set_max_active(self._h, int(queries)), "gem_hip_set_max_active")

    def set_graph(self, on=True):
        """Replay the batched path's extend step as a captured hipGraph."""
        self._check(self._lib.gem_hip_set_graph(self._h, 1 if on else 0), "gem_hip_set_graph")

    def set_table_stream(self, cache_bytes, slot_bytes=0):
        """Batched paths: keep edge-table values in pinned host memory and
        stream them through a device cache of `cache_bytes` (0 = in HBM)."""
        self._check(self._lib.gem_hip_set_table_stream(self._h, int(cache_bytes), int(slot_bytes)),
                    "gem_hip_set_table_stream")

    def stream_stats(self):
        """Streamed-table statistics since the last call (then reset)."""
        st = StreamStats()
        self._check(self._lib.gem_hip_stream_stats(self._h, C.byref(st)), "gem_hip_stream_stats")
        return {k: getattr(st, k) for k, _ in StreamStats._fields_}

    def set_work_limit(self, rows):
        """Stop a join after `rows` processed rows (0 = no limit)."""
        self._check(self._lib.gem_hip_set_work_limit(self._h, int(rows)), "gem_hip_set_work_limit")

    def match_full(self, query, cap=0):
        """Whole plain pipeline with GEM order: dict(count, saturated,
        work_limited, rows_processed, order)."""
        self.query = query
        res = JoinResult()
        order = np.zeros(query.num_e, np.uint32)
        v = query.view()
        self._check(self._lib.gem_hip_match(self._h, C.byref(v), cap, C.byref(res), _p32(order)), "gem_hip_match")
        return {"count": int(res.count), "saturated": bool(res.saturated), "work_limited": bool(res.work_limited),
                "rows_processed": int(res.rows_processed), "work": int(res.work), "order": order}

    def vjoin(self, vertex_order=None, cap=0):
        """Vertex-at-a-time join of the current query (after filter and
        build_edge_tables); vertex_order None = from GEM's edge order."""
        res = JoinResult()
        lr = np.zeros(GEM_MAX_QV, np.uint64)
        vo = None if vertex_order is None else np.ascontiguousarray(vertex_order, dtype=np.uint32)
        self._check(self._lib.gem_hip_vjoin(self._h, _p32(vo) if vo is not None else None, cap, C.byref(res),
                                            _p64(lr)), "gem_hip_vjoin")
        return {"count": int(res.count), "saturated": bool(res.saturated), "work_limited": bool(res.work_limited),
                "rows_processed": int(res.rows_processed), "work": int(res.work),
                "level_rows": lr[:self.query.num_v].copy()}

    def match_batch(self, queries, vertex_orders=None, cap=0, work_limit=0, level_budget=0, group=0):
        """Batched vertex-at-a-time path over a list of queries.  Returns a
        list of dict(count, saturated, work_limited, rows_processed, level_rows)."""
        nq = len(queries)
        views = (QueryView * max(nq, 1))(*[q.view() for q in queries])
        keep = []
        vo_arr = None
        if vertex_orders is not None:
            vo_arr = (C.POINTER(C.c_uint32) * max(nq, 1))()
            for i, vo in enumerate(vertex_orders):
                if vo is None:
                    vo_arr[i] = C.POINTER(C.c_uint32)()
                else:
                    a = np.ascontiguousarray(vo, dtype=np.uint32)
                    keep.append(a)
                    vo_arr[i] = _p32(a)
        opts = BatchOpts(int(cap), int(work_limit), int(level_budget), int(group), 0, 0, 0, 0)
        res = (JoinResult * max(nq, 1))()
        lr = np.zeros(max(nq, 1) * GEM_MAX_QV, np.uint64)
        self._check(self._lib.gem_hip_match_batch(self._h, views, nq, vo_arr, C.byref(opts), res, _p64(lr)),
                    "gem_hip_match_batch")
        out = []
        for i, q in enumerate(queries):
            r = res[i]
            out.append({"count": int(r.count), "saturated": bool(r.saturated), "work_limited": bool(r.work_limited),
                        "rows_processed": int(r.rows_processed), "work": int(r.work),
                        "level_rows": lr[i * GEM_MAX_QV:i * GEM_MAX_QV + q.num_v].copy()})
        return out

    def match_relaxed(self, query, threshold, model=None, max_orders=8, cap=0, work_limit=0, level_budget=0,
                      group=0, first_attempt=0, restart_growth=4, restart_width=1, max_attempts=0):
        """Relaxed matching (gem_hip_match_relaxed): one entry per relaxed
        query of `query` at `threshold`, in the library's order, as
        dict(mask, count, saturated, work_limited, work, restarts); mask is an
        int with bit i set when edge i is kept."""
        from .graph import relaxed_queries, masks_to_ints
        n = len(relaxed_queries(query, threshold))
        v = query.view()
        opts = BatchOpts(int(cap), int(work_limit), int(level_budget), int(group), int(restart_growth),
                         int(first_attempt), int(restart_width), int(max_attempts))
        masks = np.zeros((n, GEM_RQ_WORDS), np.uint64)
        res = (JoinResult * n)()
        got = C.c_uint32()
        self._check(self._lib.gem_hip_match_relaxed(self._h, C.byref(v), int(threshold),
                                                    model._h if model is not None else None, int(max_orders),
                                                    C.byref(opts), n, _p64(masks), res, C.byref(got)),
                    "gem_hip_match_relaxed")
        return [{"mask": m, "count": int(r.count), "saturated": bool(r.saturated),
                 "work_limited": bool(r.work_limited), "work": int(r.work), "restarts": int(r.restarts)}
                for m, r in zip(masks_to_ints(masks[:got.value]), res)]

    def export_embeddings(self, query, path, max_rows=0, run_bytes=0):
        """Every embedding of `query` (plain pipeline) in a canonical file
        (gem_amd.export); dict(rows, digest, runs, num_v)."""
        from ._lib import ExportResult
        r = ExportResult()
        v = query.view()
        self._check(self._lib.gem_hip_export_embeddings(self._h, C.byref(v), str(path).encode(), int(max_rows),
                                                        int(run_bytes), C.byref(r)), "gem_hip_export_embeddings")
        return {"rows": int(r.rows), "digest": int(r.digest), "runs": int(r.runs), "num_v": int(r.num_v)}

    def match(self, query, cap=0):
        """Whole plain pipeline with GEM order; returns (count, saturated, order)."""
        r = self.match_full(query, cap)
        if r["work_limited"]:
            raise GemError(7, "join stopped at the work limit")
        return r["count"], r["saturated"], r["order"]
