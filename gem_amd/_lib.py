"""ctypes binding of gem_amd/_lib/libgem_hip.so (the C ABI in include/gem_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C gem_amd/csrc``).  Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libgem_hip.so")

GEM_MAX_QV = 32
GEM_MAX_QE = 256

STATUS = {0: "GEM_OK", 1: "GEM_E_INVALID", 2: "GEM_E_HIP", 3: "GEM_E_NOMEM", 4: "GEM_E_QUERY",
          5: "GEM_E_STATE", 6: "GEM_E_IO", 7: "GEM_E_LIMIT"}

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class QueryView(C.Structure):
    _fields_ = [("num_v", C.c_uint32), ("num_e", C.c_uint32), ("labels", u32p), ("edges", u32p)]


class CsrView(C.Structure):
    _fields_ = [("num_v", C.c_uint32), ("num_slots", C.c_uint64), ("offsets", u64p),
                ("neighbors", u32p), ("labels", u32p)]


class JoinResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("saturated", C.c_uint32), ("num_steps", C.c_uint32),
                ("num_embeddings_written", C.c_uint64), ("rows_processed", C.c_uint64),
                ("work_limited", C.c_uint32), ("restarts", C.c_uint32), ("work", C.c_uint64)]


class StepLog(C.Structure):
    _fields_ = [("edge", C.c_uint32 * GEM_MAX_QE), ("is_select", C.c_uint32 * GEM_MAX_QE),
                ("rows", C.c_uint64 * GEM_MAX_QE), ("ms", C.c_float * GEM_MAX_QE)]


class BatchOpts(C.Structure):
    _fields_ = [("cap", C.c_uint64), ("work_limit", C.c_uint64), ("level_budget", C.c_uint64),
                ("group", C.c_uint32), ("restart_growth", C.c_uint32), ("first_attempt", C.c_uint64),
                ("restart_width", C.c_uint32), ("max_attempts", C.c_uint32)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("launches", C.c_uint64), ("ms", C.c_double), ("bytes", C.c_double)]


class StreamStats(C.Structure):
    _fields_ = [("table_bytes", C.c_uint64), ("h2d_bytes", C.c_uint64), ("blocks_loaded", C.c_uint64),
                ("blocks_hit", C.c_uint64), ("groups", C.c_uint64), ("serial_groups", C.c_uint64),
                ("copy_ms", C.c_double), ("step_ms", C.c_double)]


class MemoryStats(C.Structure):
    _fields_ = [("in_use", C.c_uint64), ("peak_in_use", C.c_uint64), ("held", C.c_uint64), ("reusable", C.c_uint64),
                ("level_budget", C.c_uint64), ("peak_charged", C.c_uint64), ("iterations", C.c_uint64),
                ("forced", C.c_uint64)]


class PhaseTimes(C.Structure):
    _fields_ = [("filter_ms", C.c_float), ("index_ms", C.c_float), ("tables_ms", C.c_float),
                ("join_ms", C.c_float)]


# (name, restype, argtypes) for every function declared in include/gem_hip.h
SIGNATURES = [
    ("gem_graph_generate_er", C.c_int, [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("gem_graph_generate_rmat", C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("gem_graph_from_csr", C.c_int, [C.c_uint32, u64p, u32p, u32p, C.POINTER(C.c_void_p)]),
    ("gem_graph_from_csr32", C.c_int, [C.c_uint32, u32p, u32p, u32p, C.POINTER(C.c_void_p)]),
    ("gem_graph_load", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("gem_graph_save", C.c_int, [C.c_void_p, C.c_char_p]),
    ("gem_graph_save_binary", C.c_int, [C.c_void_p, C.c_char_p]),
    ("gem_graph_digest", C.c_uint64, [C.c_void_p]),
    ("gem_graph_num_vertices", C.c_uint32, [C.c_void_p]),
    ("gem_graph_num_slots", C.c_uint64, [C.c_void_p]),
    ("gem_graph_offsets", u64p, [C.c_void_p]),
    ("gem_graph_neighbors", u32p, [C.c_void_p]),
    ("gem_graph_labels", u32p, [C.c_void_p]),
    ("gem_graph_free", None, [C.c_void_p]),
    ("gem_query_random_walk", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, u32p, u32p, C.c_uint32, u32p]),
    ("gem_query_load", C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p]),
    ("gem_query_save", C.c_int, [C.POINTER(QueryView), C.c_char_p]),
    ("gem_plan_gem_order", C.c_int, [C.POINTER(QueryView), u64p, u32p]),
    ("gem_plan_relaxed_queries", C.c_int, [C.POINTER(QueryView), C.c_uint32, C.c_uint32, u64p, u32p]),
    ("gem_hip_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("gem_hip_destroy", None, [C.c_void_p]),
    ("gem_hip_last_error", C.c_char_p, [C.c_void_p]),
    ("gem_hip_set_level_budget", C.c_int, [C.c_void_p, C.c_uint64]),
    ("gem_hip_set_work_limit", C.c_int, [C.c_void_p, C.c_uint64]),
    ("gem_hip_set_persistent_blocks", C.c_int, [C.c_void_p, C.c_uint32]),
    ("gem_hip_set_heavy_split", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    ("gem_hip_set_graph", C.c_int, [C.c_void_p, C.c_int]),
    ("gem_hip_set_graph_residency", C.c_int, [C.c_void_p, C.c_int]),
    ("gem_hip_set_relaxed_sharing", C.c_int, [C.c_void_p, C.c_int]),
    ("gem_hip_set_racing", C.c_int, [C.c_void_p, C.c_int]),
    ("gem_hip_set_prep_overlap", C.c_int, [C.c_void_p, C.c_int]),
    ("gem_hip_set_level_swap", C.c_int, [C.c_void_p, C.c_uint64]),
    ("gem_hip_swap_stats", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gem_export_rows", C.c_int, [u32p, C.c_uint64, C.c_uint32, C.c_char_p, C.c_uint64, C.c_void_p]),
    ("gem_hip_export_embeddings", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_char_p, C.c_uint64, C.c_uint64,
                                            C.c_void_p]),
    ("gem_hip_set_max_active", C.c_int, [C.c_void_p, C.c_uint32]),
    ("gem_hip_set_table_stream", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    ("gem_hip_stream_stats", C.c_int, [C.c_void_p, C.POINTER(StreamStats)]),
    ("gem_hip_load_data_csr", C.c_int, [C.c_void_p, C.POINTER(CsrView)]),
    ("gem_hip_filter", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_uint32, u32p, u32p, C.c_uint64]),
    ("gem_hip_filter_noalloc", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_uint64]),
    ("gem_hip_position_lookup", C.c_int, [C.c_void_p, C.c_uint32, u32p, C.c_uint64, u32p]),
    ("gem_hip_build_edge_tables", C.c_int, [C.c_void_p, u64p]),
    ("gem_hip_read_edge_table", C.c_int, [C.c_void_p, C.c_uint32, u64p, u32p]),
    ("gem_hip_join", C.c_int, [C.c_void_p, u32p, C.c_uint64, C.POINTER(JoinResult), C.POINTER(StepLog),
                               u32p, C.c_uint64]),
    ("gem_hip_phase_times", C.c_int, [C.c_void_p, C.POINTER(PhaseTimes)]),
    ("gem_hip_memory_stats", C.c_int, [C.c_void_p, C.c_int, C.POINTER(MemoryStats)]),
    ("gem_hip_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("gem_hip_kernel_stats", C.c_int, [C.c_void_p, C.POINTER(KernelStat), C.c_uint32, u32p]),
    ("gem_hip_match", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_uint64, C.POINTER(JoinResult), u32p]),
    ("gem_hip_match_batch", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_uint32, C.POINTER(u32p),
                                      C.POINTER(BatchOpts), C.POINTER(JoinResult), u64p]),
    ("gem_hip_vjoin", C.c_int, [C.c_void_p, u32p, C.c_uint64, C.POINTER(JoinResult), u64p]),
    ("gem_mlp_create", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("gem_mlp_destroy", None, [C.c_void_p]),
    ("gem_mlp_num_params", C.c_uint64, [C.c_void_p]),
    ("gem_mlp_get_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("gem_mlp_set_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("gem_mlp_set_margin", C.c_int, [C.c_void_p, C.c_float]),
    ("gem_mlp_predict", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_uint32, C.POINTER(C.c_float)]),
    ("gem_mlp_loss_grad", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_uint32,
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("gem_mlp_train", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_uint32, C.c_uint32,
                                C.c_float, C.POINTER(C.c_float)]),
    ("gem_hip_order_candidates", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, u32p, C.POINTER(C.c_int32), u32p]),
    ("gem_hip_match_batch_learned", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_uint32, C.c_void_p,
                                              C.c_uint32, C.POINTER(BatchOpts), C.POINTER(JoinResult), u64p, u32p]),
    ("gem_hip_match_relaxed", C.c_int, [C.c_void_p, C.POINTER(QueryView), C.c_uint32, C.c_void_p, C.c_uint32,
                                        C.POINTER(BatchOpts), C.c_uint32, u64p, C.POINTER(JoinResult), u32p]),
]

class SwapStats(C.Structure):
    _fields_ = [("levels_out", C.c_uint64), ("levels_in", C.c_uint64), ("bytes_out", C.c_uint64),
                ("bytes_in", C.c_uint64), ("peak_host_bytes", C.c_uint64)]


class ExportResult(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("digest", C.c_uint64), ("runs", C.c_uint32), ("num_v", C.c_uint32)]


GEM_RQ_WORDS = 4
GEM_MAX_RQ = 1 << 20

GEM_FEAT_DIM = 272

_lib = None


def load():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           f"or make -C gem_amd/csrc")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GemError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status
