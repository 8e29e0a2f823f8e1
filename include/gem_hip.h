/*
 * gem_hip.h — C ABI of the MI355X (gfx950) subgraph-matching library.
 *
 * The library enumerates/counts the embeddings of a labelled query graph in a
 * labelled data graph (injective, label- and edge-preserving maps), and of its
 * relaxed queries at a threshold T > 0, with GEM's edge-centric pipeline:
 *
 *   data CSR -> candidate filter -> position index -> edge-candidate tables
 *            -> GEM-ordered join (expand / select) -> uint64 count
 *
 * Every entry point returns an int status (GEM_OK == 0) and never exits the
 * process; the message of the last failure on a context is available from
 * gem_hip_last_error().  Host arrays are caller-owned and are copied by the
 * library; device memory is owned by the opaque gem_ctx.  One context per host
 * thread.  No torch or HIP types appear in these signatures.
 *
 * Reference interfaces replaced (paths relative to the Seungsu-J/GEM checkout):
 *   in-core/lib/io/include/io.h:7-27            readGraphToCPUStandard / copyGraphToGPU
 *   in-core/lib/filter/include/filter.h:10-14   filterSSM
 *   in-core/lib/structure/include/unifiedTrie.cuh:58-61,244-251
 *                                               HashTable::build, initVertexTries,
 *                                               construct_edge_candidates(_cpu)
 *   in-core/lib/structure/src/unifiedTrie.cu:939-1018
 *                                               find_best_edge_candidate (GEM order)
 *   in-core/lib/join/include/join_trie.cuh:31-32 join
 *   in-core/lib/plan/include/relax.h            relax (relaxed queries, T > 0)
 *   in-core/src/main.cu:229-236                 per-query result readout
 *   out-of-core/src/main.cu:334-410             per-query child processes (CLI)
 */
#ifndef GEM_HIP_H
#define GEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum gem_status {
  GEM_OK = 0,
  GEM_E_INVALID = 1, /* bad argument / malformed input                      */
  GEM_E_HIP = 2,     /* HIP runtime error (message names the call)          */
  GEM_E_NOMEM = 3,   /* device or host allocation failed                    */
  GEM_E_QUERY = 4,   /* query unsupported: disconnected, < 2 edges, too big */
  GEM_E_STATE = 5,   /* call out of order (e.g. join before filter)         */
  GEM_E_IO = 6,      /* file could not be read or written                   */
  GEM_E_LIMIT = 7    /* a documented size limit was exceeded                */
};

/* Query limits.  The reference caps queries at 16 vertices / 16-32 edges
 * (in-core/lib/utils/include/defs.h:20-23); these limits are wider. */
#define GEM_MAX_QV 32u
#define GEM_MAX_QE 256u

/* ------------------------------------------------------- host graph (CSR) */
/* A simple undirected labelled graph held in host memory as CSR with sorted,
 * duplicate-free adjacency and u64 offsets.  Replaces cpuGraph
 * (in-core/lib/graph/include/cpuGraph.h:12-67) on the data side. */
typedef struct gem_graph gem_graph;

/* Erdos-Renyi G(n, m): m distinct unordered pairs, labels uniform in
 * [0, num_labels).  Deterministic in seed. */
int gem_graph_generate_er(uint32_t n, uint64_t m, uint32_t num_labels,
                          uint64_t seed, gem_graph **out);
/* R-MAT with 2^scale vertices and round(2^scale * avg_degree / 2) samples,
 * quadrant probabilities (a, b, c, 1-a-b-c); symmetrised, self-loops and
 * duplicates dropped.  label_dist: 0 = uniform, 1 = Zipf (s = 1). */
int gem_graph_generate_rmat(uint32_t scale, double avg_degree, double a,
                            double b, double c, uint32_t num_labels,
                            uint32_t label_dist, uint64_t seed, gem_graph **out);
/* Copy an existing CSR (offsets[n+1], neighbors[offsets[n]], labels[n]).
 * Adjacency must be sorted ascending, without self-loops or duplicates, and
 * symmetric; GEM_E_INVALID otherwise. */
int gem_graph_from_csr(uint32_t n, const uint64_t *offsets,
                       const uint32_t *neighbors, const uint32_t *labels,
                       gem_graph **out);
/* The reference's host CSR as its loader leaves it: cpuGraph::offsets_ (u32,
 * [n+1]), neighbors_ (file order per vertex) and vLabels_
 * (in-core/lib/graph/include/cpuGraph.h:23-31, filled by
 * readGraphToCPUStandard, in-core/lib/io/src/io.cu:69-199).  Offsets are
 * widened, each adjacency list is sorted; self-loops, repeated neighbours,
 * asymmetric edges or neighbours out of range give GEM_E_INVALID.  Labels are
 * taken as given (see the integration notes for the reference's per-query
 * label ids). */
int gem_graph_from_csr32(uint32_t n, const uint32_t *offsets,
                         const uint32_t *neighbors, const uint32_t *labels,
                         gem_graph **out);
/* Reference text format: "t N M" / "v id label degree" / "e src dst"
 * (in-core/lib/io/src/io.cu:45-220).  Declared degrees must match. */
int gem_graph_load(const char *path, gem_graph **out);
int gem_graph_save(const gem_graph *g, const char *path);
/* Binary CSR cache (no reference counterpart: the reference parses the text
 * format on every run, in-core/lib/io/src/io.cu:45-220, which takes minutes
 * for graphs of billions of edges).  A 64-byte header ("GEMCSR01", n, slots,
 * a digest of the arrays), then offsets, neighbors and labels as in memory.
 * gem_graph_load recognises such a file by its magic and checks the digest,
 * the offsets and the neighbour range (GEM_E_INVALID otherwise). */
int gem_graph_save_binary(const gem_graph *g, const char *path);
/* FNV-1a 64 digest of the CSR arrays (over 1 MiB pieces, combined in order). */
uint64_t gem_graph_digest(const gem_graph *g);
uint32_t gem_graph_num_vertices(const gem_graph *g);
uint64_t gem_graph_num_slots(const gem_graph *g); /* = 2 |E| */
const uint64_t *gem_graph_offsets(const gem_graph *g);
const uint32_t *gem_graph_neighbors(const gem_graph *g);
const uint32_t *gem_graph_labels(const gem_graph *g);
void gem_graph_free(gem_graph *g);

/* ------------------------------------------------------------------- query */
/* A query: num_v vertices with ORIGINAL labels, num_e undirected edges given as
 * (src, dst) pairs in file order with src < dst.  Edge i is the reference's
 * edge-tag bit i; directed edge 2i is src->dst and 2i+1 is dst->src
 * (in-core/lib/io/src/io.cu:176-199). */
typedef struct gem_query_view {
  uint32_t num_v;
  uint32_t num_e;
  const uint32_t *labels; /* [num_v] */
  const uint32_t *edges;  /* [2 * num_e] */
} gem_query_view;

/* Random-walk query: walk from a seeded start until k distinct vertices are
 * visited; the query is the induced subgraph on them (vertex ids in visit
 * order, labels copied, edges sorted).  Returns GEM_E_LIMIT when the induced
 * subgraph has more than max_e edges. */
int gem_query_random_walk(const gem_graph *g, uint32_t k, uint64_t seed,
                          uint32_t *labels_out, uint32_t *edges_out,
                          uint32_t max_e, uint32_t *num_e_out);
int gem_query_load(const char *path, uint32_t max_v, uint32_t max_e,
                   uint32_t *num_v, uint32_t *labels, uint32_t *num_e,
                   uint32_t *edges);
int gem_query_save(const gem_query_view *q, const char *path);

/* ------------------------------------------------------------------- plan */
/* GEM ordering at threshold 0 (in-core/lib/structure/src/unifiedTrie.cu:939-1018
 * with the lattice of in-core/lib/structure/src/lattice.cpp:41-185, 322-467):
 * first edge = argmax 0.6/(2+|T_e|) + 0.3c + 0.01 P_e (strict >, edge index
 * order); then repeatedly the adjacent edge maximising 0.4/(2+|T_e|) + 0.4c,
 * ties broken by the reference's lattice-id order.  table_sizes[i] = |T_e| of
 * undirected edge i.  order_out[num_e] receives edge indices. */
int gem_plan_gem_order(const gem_query_view *q, const uint64_t *table_sizes,
                       uint32_t *order_out);

/* Relaxed queries (relax(), in-core/lib/plan/src/relax.cpp:84-142, called at
 * in-core/src/main.cu:88-92): the edge subsets of q that leave out at most
 * `threshold` edges and still connect every query vertex (the reference's
 * checkConnectivity, relax.cpp:57-81).  Threshold 0 gives the full query
 * only.  Order: the full set, then by the number of edges left out, then by
 * the ascending tuple of left-out edge indices.  Edge set i is
 * masks[i * GEM_RQ_WORDS .. +GEM_RQ_WORDS), bit b of word w = edge 64 w + b
 * (the reference's ettype bitset, in-core/lib/utils/include/defs.h:23-39,
 * holds 32 edges; this one holds GEM_MAX_QE).  max_rq = 0 only counts
 * (*num_rq); more than max_rq sets (or GEM_MAX_RQ when counting) is
 * GEM_E_LIMIT. */
#define GEM_RQ_WORDS (GEM_MAX_QE / 64u)
#define GEM_MAX_RQ (1u << 20)
int gem_plan_relaxed_queries(const gem_query_view *q, uint32_t threshold, uint32_t max_rq,
                             uint64_t *masks, uint32_t *num_rq);

/* -------------------------------------------------------- device context */
/* A context owns one HIP stream on `device`; all of its device work runs on
 * that stream.  Device memory the library frees is kept for reuse by the same
 * context (large buffers are carved from 8 GiB segments) and is returned to
 * the driver when the context is destroyed, or earlier when the driver refuses
 * an allocation. */
typedef struct gem_ctx gem_ctx;

int gem_hip_create(int device, gem_ctx **out);
void gem_hip_destroy(gem_ctx *ctx);
const char *gem_hip_last_error(const gem_ctx *ctx);
/* Largest join level materialised at once (bytes, default 2 GiB, minimum
 * 1 MiB); a larger level is split into row ranges finished depth-first.  No
 * reference counterpart: the reference materialises every level whole and
 * relies on its memory manager (in-core/lib/join/src/join_trie.cu:1087-1277). */
int gem_hip_set_level_budget(gem_ctx *ctx, uint64_t bytes);

/* Data graph view (plain host pointers) uploaded once per context.
 * Replaces allocateMemGPU/copyGraphToGPU (in-core/lib/io/src/io.cu:348-452). */
typedef struct gem_csr_view {
  uint32_t num_v;
  uint64_t num_slots;
  const uint64_t *offsets;   /* [num_v + 1] */
  const uint32_t *neighbors; /* [num_slots], sorted per vertex */
  const uint32_t *labels;    /* [num_v], original labels */
} gem_csr_view;
int gem_hip_load_data_csr(gem_ctx *ctx, const gem_csr_view *g);
/* Where the next gem_hip_load_data_csr keeps the neighbour array (the
 * out-of-core reference keeps its data graph on the host, out-of-core/src/
 * main.cu:445-447, and filters and builds tables on the CPU,
 * out-of-core/lib/filter/src/filter.cu:61-295 and
 * out-of-core/lib/structure/src/unifiedTrie.cu:1224-1427): host = 0, in HBM
 * (default); host = 1, in pinned, mapped host memory that the filter and
 * table kernels read over PCIe, for graphs larger than HBM.  Offsets and
 * labels (12 B per vertex) stay in HBM.  Results do not depend on it. */
int gem_hip_set_graph_residency(gem_ctx *ctx, int host);

/* Candidate filter for a new query (replaces filterSSM,
 * in-core/lib/filter/src/filter.cu:34-167 and the CPU form
 * out-of-core/lib/filter/src/filter.cu:61-295):
 *   Can(u) = { v asc : lab(v) = lab(u), deg(v) + T >= deg(u),
 *              sum_l max(0, nlc_u[l] - nlc_v[l]) <= T }.
 * Makes `q` the context's current query.  h_num_can[num_v] receives |Can(u)|;
 * if h_can is non-null, row u of h_can (pitch can_pitch) receives Can(u). */
int gem_hip_filter(gem_ctx *ctx, const gem_query_view *q, uint32_t threshold,
                   uint32_t *h_num_can, uint32_t *h_can, uint64_t can_pitch);

/* gem_hip_filter without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Outputs go
 * to caller-owned DEVICE arrays: d_num_can[num_v] and, when d_can is
 * non-null, Can(u) ascending at d_can + u * can_pitch (can_pitch must be at
 * least the data frequency of every query vertex's label).  The filter's
 * intermediate buffers live in one fixed 64 MiB device array per device
 * (held until the process ends; its address is found when a context is
 * created): a graph and query needing more (about 12 B per data vertex per
 * query vertex plus 4 B per data vertex) fail with GEM_E_LIMIT.  Calls that
 * overlap on one device share that array and race.  The context's current
 * query (gem_hip_filter's state) is not touched.  Returns when the outputs are
 * written. */
int gem_hip_filter_noalloc(gem_ctx *ctx, const gem_query_view *q, uint32_t threshold,
                           uint32_t *d_num_can, uint32_t *d_can, uint64_t can_pitch);

/* Position index probe (the result the reference's cuckoo hash returns,
 * in-core/lib/structure/include/unifiedTrie.cuh:84-125): pos_out[i] =
 * index of vertices[i] in Can(u), or 0xFFFFFFFF when it is not a candidate. */
int gem_hip_position_lookup(gem_ctx *ctx, uint32_t u, const uint32_t *vertices,
                            uint64_t n, uint32_t *pos_out);

/* Edge-candidate tables of the current query (replaces
 * construct_edge_candidates, in-core/lib/structure/src/unifiedTrie.cu:1020-1391,
 * and construct_edge_candidates_cpu, out-of-core/.../unifiedTrie.cu:1224-1427):
 * for every directed query edge d = u->u' with deg_q(u) != 1, row i holds the
 * neighbours w of Can(u)[i] that lie in Can(u'), in adjacency order, stored as
 * their positions in Can(u').  h_table_sizes[2*num_e] receives |T_d| (0 for
 * directions keyed on a leaf, which have no table). */
int gem_hip_build_edge_tables(gem_ctx *ctx, uint64_t *h_table_sizes);

/* Copy table d back: offsets[|Can(u)|+1] (u64), values[|T_d|] (positions in
 * Can(u')).  For parity checks. */
int gem_hip_read_edge_table(gem_ctx *ctx, uint32_t directed_edge,
                            uint64_t *offsets, uint32_t *values);

/* Result of one query. */
typedef struct gem_join_result {
  uint64_t count;      /* embeddings (exact when saturated == 0)         */
  uint32_t saturated;  /* 1 when the count exceeded `cap` (count == cap) */
  uint32_t num_steps;  /* join steps run (= num_e - 1)                   */
  uint64_t num_embeddings_written; /* rows stored to emb_out            */
  uint64_t rows_processed; /* rows read by the join's steps              */
  uint32_t work_limited;   /* 1: stopped at the work limit, count unknown */
  uint32_t restarts;       /* batched learned path: attempts that went past
                              their work limit and were started again with
                              the next ranked order (0 elsewhere)          */
  uint64_t work;           /* what the work limit bounds: rows read (plain
                              path); rows read + list entries walked, upper
                              bound (batched path)                         */
} gem_join_result;

/* Stop a join once it has processed more than `rows` rows (0 = no limit);
 * the result then has work_limited = 1 and no valid count.  No reference
 * counterpart (the reference has no time or work limit). */
int gem_hip_set_work_limit(gem_ctx *ctx, uint64_t rows);

/* Optional per-step log: rows after step s (s = 0 is the first edge table),
 * time of the step in milliseconds, and whether it was a select. */
typedef struct gem_step_log {
  uint32_t edge[GEM_MAX_QE];
  uint32_t is_select[GEM_MAX_QE];
  uint64_t rows[GEM_MAX_QE];
  float ms[GEM_MAX_QE];
} gem_step_log;

/* Join (replaces join(), in-core/lib/join/src/join_trie.cu:1087-1277) of the
 * current query along edge_order[num_e] (any connected order; GEM order from
 * gem_plan_gem_order reproduces the reference's plan).  The count stops at
 * `cap` (0 = no cap): saturated = 1 iff the true count exceeds cap.  If
 * emb_out is non-null and the count is <= emb_capacity, the embeddings are
 * written as rows of num_v vertex ids (column u = image of query vertex u),
 * in unspecified order. */
int gem_hip_join(gem_ctx *ctx, const uint32_t *edge_order, uint64_t cap,
                 gem_join_result *out, gem_step_log *log, uint32_t *emb_out,
                 uint64_t emb_capacity);

/* Phase timings (milliseconds, HIP events on the context stream) of the
 * last filter / table build / join, mirroring the reference's printed phases
 * (in-core/src/main.cu:147-227). */
typedef struct gem_phase_times {
  float filter_ms;
  float index_ms;
  float tables_ms;
  float join_ms;
} gem_phase_times;
int gem_hip_phase_times(const gem_ctx *ctx, gem_phase_times *out);

/* Kernel-level profiling (no reference counterpart; the reference times
 * phases on the host only, in-core/src/main.cu:147-227).  While on, every
 * launch of the library's kernels is bracketed by HIP events on the context
 * stream and its algorithmic bytes (compulsory traffic) are accumulated.
 * gem_hip_set_profiling(ctx, 1) also clears the statistics. */
typedef struct gem_kernel_stat {
  char name[40];
  uint64_t launches;
  double ms;    /* summed launch durations */
  double bytes; /* summed algorithmic bytes */
} gem_kernel_stat;
int gem_hip_set_profiling(gem_ctx *ctx, int on);
int gem_hip_kernel_stats(gem_ctx *ctx, gem_kernel_stat *out, uint32_t max, uint32_t *n);

/* Device memory of the library's caching allocator (shared by every context
 * of the process) and the memory figures of this context's last batched run.
 * No reference counterpart: the reference's MemoryManager
 * (out-of-core/lib/structure/src/memory_manager.cu:196-822) reports none. */
typedef struct gem_memory_stats {
  uint64_t in_use;       /* bytes of live device buffers                      */
  uint64_t peak_in_use;  /* highest in_use since the last reset               */
  uint64_t held;         /* bytes obtained from the driver                    */
  uint64_t reusable;     /* held but unused: empty segments, parked blocks    */
  uint64_t level_budget; /* last batched run: bytes its pending level rows
                            were charged against                             */
  uint64_t peak_charged; /* last batched run: most pending + newly allocated
                            level bytes at once                              */
  uint64_t iterations;   /* last batched run: scheduler iterations            */
  uint64_t forced;       /* last batched run: iterations that went past the
                            budget because nothing else could move           */
} gem_memory_stats;
/* reset_peak != 0: restart peak_in_use from the current in_use afterwards. */
int gem_hip_memory_stats(gem_ctx *ctx, int reset_peak, gem_memory_stats *out);

/* ------------------------------------------------------- batched path */
/* The batched path has no single reference counterpart: it is the
 * reference's per-query pipeline (in-core/src/main.cu:148-236, join() at
 * in-core/lib/join/src/join_trie.cu:1087-1277) run for many queries at once,
 * with the join done vertex-at-a-time: every edge between a new query vertex
 * and earlier ones is applied in the same step (intersection of table rows).
 * Counts equal the plain path's (the count does not depend on the order). */
typedef struct gem_batch_opts {
  uint64_t cap;          /* per-query count cap (0 = none)                  */
  uint64_t work_limit;   /* work a query may do before it stops unfinished:
                            rows read plus list entries walked, summed over
                            its attempts (0 = none)                        */
  uint64_t level_budget; /* bytes for pending level chunks, shared by the
                            group's running queries (0 = a quarter of the
                            device memory free when the join starts)       */
  uint32_t group;        /* queries resident on the GPU at once (0 = all)   */
  uint32_t restart_growth; /* learned path: factor between the work limits
                            of successive rounds of attempts (0 = 4)       */
  uint64_t first_attempt; /* learned path: work limit of the first round's
                            attempts (0 = no restarts).  Attempt a runs the
                            a-th ranked order (once all n have run: the one
                            whose attempts counted the most embeddings per
                            unit of work) and may do
                            first_attempt * growth^(a / width) work before
                            the query starts again                        */
  uint32_t restart_width; /* learned path: attempts per round (0 = 1)       */
  uint32_t max_attempts;  /* learned path: attempts per query (0 = no bound;
                            otherwise the last one has no limit)           */
} gem_batch_opts;

/* Match nq queries.  vertex_orders (nullable, and each entry nullable) gives a
 * query's vertex order (vertex_orders[i][0] must not be a leaf, every later
 * vertex needs an earlier neighbour); absent orders come from GEM's edge order
 * (first appearance).  results[nq]; level_rows (nullable) receives, per query,
 * GEM_MAX_QV entries: rows after each vertex level. */
int gem_hip_match_batch(gem_ctx *ctx, const gem_query_view *queries, uint32_t nq,
                        const uint32_t *const *vertex_orders, const gem_batch_opts *opts,
                        gem_join_result *results, uint64_t *level_rows);

/* Grid size (workgroups of 256 threads) of the batched path's persistent
 * kernels; they pull tasks from a device counter, so this sets occupancy,
 * not the amount of work (default 2048).  Set by the autotuner. */
int gem_hip_set_persistent_blocks(gem_ctx *ctx, uint32_t blocks);
/* Load balance of the batched path's walks: a row whose driving (shortest)
 * list has more than `threshold` entries is set aside and walked `chunk`
 * entries per wave task by a second kernel instead of by the wave holding the
 * row (defaults 1024 and 1024; threshold in [32, 2^24], chunk in
 * [64, 2^20]).  Counts do not depend on either.  Set by the autotuner. */
int gem_hip_set_heavy_split(gem_ctx *ctx, uint32_t threshold, uint32_t chunk);
/* on != 0: the batched path's per-iteration extend step (table uploads, three
 * persistent kernels, result download) is captured once per run as a hipGraph
 * and replayed; off (default): launched eagerly.  Kernel profiling forces the
 * eager form. */
int gem_hip_set_graph(gem_ctx *ctx, int on);
/* Queries of a batched run that hold join levels at once (default 64; 0 =
 * all): the others wait and start, in order, as running ones end (continuous
 * admission); each running query's pending rows are held to
 * level budget / this many bytes while queries wait, and to an equal share of
 * the budget over the running queries after that.  Counts do not depend on
 * it.  Set by the autotuner. */
int gem_hip_set_max_active(gem_ctx *ctx, uint32_t queries);
/* on != 0 (default): in the learned path with restarts, every query with at
 * least two candidate orders gets a twin that runs another order; twins start
 * once every query of the batch has started, as running queries end (the
 * batch's tail), at their query's current attempt limit, and whichever of a
 * pair ends first with a count cancels the other.  Results are the same; a
 * query's `work` includes its twin's.  Not used with streamed tables. */
int gem_hip_set_racing(gem_ctx *ctx, int on);
/* on != 0 (default): in the batched learned path without streamed tables,
 * with continuous admission (max_active != 0) and profiling off, the first
 * max_active + 16 queries of a group are filtered and tabled before the join
 * starts and the rest by a second context of this one (its own stream, made at
 * the first such run, destroyed with this one) in a host thread while the join
 * runs; a query starts once it is prepared, or the join waits for it when
 * nothing else runs.  Results are the same. */
int gem_hip_set_prep_overlap(gem_ctx *ctx, int on);
/* Swapping of pending join levels (the reference's out-of-core LRU swap of
 * intermediate results, out-of-core/lib/structure/src/memory_manager.cu:748-822
 * with LRUEvictPolicy, policy.cpp:106-172, hooked into the join at
 * out-of-core/lib/join/src/join_trie.cu:1088-1108).  host_bytes != 0: in the
 * batched paths, when the device-resident pending level rows exceed the level
 * budget after an extend step, the least recently used pending levels that are
 * not on top of a query's depth-first stack move to pinned host memory (at
 * most host_bytes in all), and come back when they are on top again.  Counts
 * do not depend on it.  0 (default): levels stay on the device. */
int gem_hip_set_level_swap(gem_ctx *ctx, uint64_t host_bytes);
typedef struct gem_swap_stats {
  uint64_t levels_out, levels_in; /* levels moved to host memory / back */
  uint64_t bytes_out, bytes_in;   /* their bytes                        */
  uint64_t peak_host_bytes;       /* most bytes in host memory at once  */
} gem_swap_stats;
/* Statistics of the last batched run. */
int gem_hip_swap_stats(const gem_ctx *ctx, gem_swap_stats *out);

/* Out-of-core edge tables for the batched paths (replaces the reference's
 * out-of-core MemoryManager / LRUEvictPolicy, out-of-core/lib/structure/src/
 * memory_manager.cu:196-822 and policy.cpp:106-172, with its join hooks at
 * out-of-core/lib/join/src/join_trie.cu:1088-1108, which swap intermediate
 * columns; here the level budget already bounds intermediates, and what is
 * streamed is the edge tables).  cache_bytes != 0: after a query's tables are
 * built, their values move to pinned host memory in key-range blocks of at
 * most slot_bytes (0 = 1 MiB); each join iteration marks the blocks its rows
 * will read, and groups of its tasks run while the blocks of the next group
 * are copied host -> device on a second stream into a device cache of
 * cache_bytes.  Counts are unchanged.  cache_bytes == 0: tables stay in HBM
 * (default).  A list longer than a slot, or a 256-row task whose blocks do
 * not fit the cache, fails with GEM_E_LIMIT; cache_bytes / slot_bytes >= 256 x
 * (the most backward edges of any query vertex) always suffices.  The
 * single-query entry points
 * (gem_hip_join, gem_hip_vjoin, gem_hip_match) always keep tables in HBM. */
int gem_hip_set_table_stream(gem_ctx *ctx, uint64_t cache_bytes, uint64_t slot_bytes);

typedef struct gem_stream_stats {
  uint64_t table_bytes;    /* table values moved to host memory             */
  uint64_t h2d_bytes;      /* block bytes copied host -> device              */
  uint64_t blocks_loaded;  /* block copies                                   */
  uint64_t blocks_hit;     /* needed blocks already resident                 */
  uint64_t groups;         /* task groups launched                           */
  uint64_t serial_groups;  /* groups whose copies could not overlap compute  */
  double copy_ms;          /* time of the copy stream's work (events)        */
  double step_ms;          /* time from first copy to last kernel (events)   */
} gem_stream_stats;
/* Statistics accumulated since the last call (then reset). */
int gem_hip_stream_stats(gem_ctx *ctx, gem_stream_stats *out);

/* Vertex-at-a-time join of the current query (after gem_hip_filter and
 * gem_hip_build_edge_tables) along vertex_order[num_v]. */
int gem_hip_vjoin(gem_ctx *ctx, const uint32_t *vertex_order, uint64_t cap, gem_join_result *out,
                  uint64_t *level_rows);

/* --------------------------------------------------- learned join order */
/* No reference counterpart: GEM orders by a fixed score
 * (in-core/lib/structure/src/unifiedTrie.cu:939-1018).  Here a small MLP
 * (in -> 128 -> 128 -> 1, ReLU, bf16 weights, fp32 accumulation), trained on
 * the GPU by the library's own kernels on join costs the matcher logs,
 * predicts the cost of candidate vertex orders; the batched path runs the
 * cheapest predicted one.  GEM's order is always a candidate, and the order
 * used when no model is given. */
#define GEM_FEAT_DIM 272u
typedef struct gem_mlp gem_mlp;
/* The model lives on ctx's device.  Destroy it before ctx; if ctx is destroyed
 * first, the model's device memory is released with it and every later call
 * on the model except gem_mlp_destroy returns GEM_E_STATE. */
int gem_mlp_create(gem_ctx *ctx, uint32_t in_dim, uint32_t hidden, uint64_t seed, gem_mlp **out);
void gem_mlp_destroy(gem_mlp *m);
uint64_t gem_mlp_num_params(const gem_mlp *m);
/* flat fp32 parameters: W1[in*h], b1[h], W2[h*h], b2[h], W3[h], b3[1] */
int gem_mlp_get_params(gem_mlp *m, float *out);
int gem_mlp_set_params(gem_mlp *m, const float *in);
/* The batched learned path runs the cheapest-predicted candidate only if its
 * prediction is below GEM's order's by more than `margin` (in the model's
 * output units; default 0): GEM's order stays the fallback when the model is
 * not confident. */
int gem_mlp_set_margin(gem_mlp *m, float margin);
int gem_mlp_predict(gem_mlp *m, const int32_t *feats, uint32_t n, float *out);
/* loss = mean((y - t)^2) / 2 over the batch; grad[num_params] (nullable) */
int gem_mlp_loss_grad(gem_mlp *m, const int32_t *feats, const float *targets, uint32_t n, float *loss,
                      float *grad);
/* full-batch Adam for `iters` steps; losses[iters] (nullable) */
int gem_mlp_train(gem_mlp *m, const int32_t *feats, const float *targets, uint32_t n, uint32_t iters, float lr,
                  float *losses);

/* For the current query (after gem_hip_filter and gem_hip_build_edge_tables):
 * up to max_orders candidate vertex orders (orders[max_orders * num_v], the
 * GEM-derived order first) and their int32 features
 * (feats[max_orders * GEM_FEAT_DIM]); *n receives the number produced.  seed
 * 0 seeds the random candidates with the query's hash, as the batched learned
 * path does. */
int gem_hip_order_candidates(gem_ctx *ctx, uint32_t max_orders, uint64_t seed, uint32_t *orders, int32_t *feats,
                             uint32_t *n);

/* Batched path with model-chosen orders: for each query up to max_orders
 * candidate orders (seeded by a hash of the query) are ranked by `mlp`'s
 * predicted cost, GEM's order first unless a candidate beats it by more than
 * the model's margin (mlp == NULL: in generation order, GEM's first), and the
 * first ranked order runs.  With opts->first_attempt != 0 a query whose
 * attempt exceeds its work limit starts again with the next ranked order
 * (rounds of restart_width attempts, each round's limit restart_growth times
 * the last), so every query ends with a count unless opts->work_limit (total
 * over its attempts) stops it.  chosen (nullable) receives, per query,
 * the candidate index of the order that produced the result (0 = GEM's
 * order). */
int gem_hip_match_batch_learned(gem_ctx *ctx, const gem_query_view *queries, uint32_t nq, gem_mlp *mlp,
                                uint32_t max_orders, const gem_batch_opts *opts, gem_join_result *results,
                                uint64_t *level_rows, uint32_t *chosen);

/* Relaxed matching of one query at `threshold` (the reference's main() at
 * -t threshold, in-core/src/main.cu:88-236: relax(), the lattice join of
 * in-core/lib/join/src/join_trie.cu:1130-1267 and the "Relaxed Query ET
 * <bits>: <count> matches" lines).  The relaxed queries are
 * gem_plan_relaxed_queries's edge sets, in its order (masks[max_rq *
 * GEM_RQ_WORDS]); results[i] is the count of relaxed query i's embeddings
 * (injective, label-preserving maps that keep the edges of set i).  They run
 * as one batch of the batched learned path (mlp, max_orders and opts as in
 * gem_hip_match_batch_learned).  By default the batch shares what the
 * reference's lattice shares below the relaxed queries: one filter of the
 * query at `threshold` (its candidates admit every relaxed query's
 * embeddings) and one edge table per query edge, read by every relaxed query
 * that keeps the edge; gem_hip_set_relaxed_sharing(ctx, 0), or streamed
 * tables, filters and tables each relaxed query on its own edges instead.
 * Counts are the same either way.  *num_rq receives the number of relaxed
 * queries; more than max_rq is GEM_E_LIMIT. */
int gem_hip_match_relaxed(gem_ctx *ctx, const gem_query_view *q, uint32_t threshold, gem_mlp *mlp,
                          uint32_t max_orders, const gem_batch_opts *opts, uint32_t max_rq, uint64_t *masks,
                          gem_join_result *results, uint32_t *num_rq);
int gem_hip_set_relaxed_sharing(gem_ctx *ctx, int on);

/* ------------------------------------------------------ embedding export */
/* No reference counterpart: the reference only counts (in-core/lib/join/src/
 * join_trie.cu:1204 never materialises final rows).  An embedding file holds
 * rows of num_v vertex ids (column u = image of query vertex u) in canonical
 * order, ascending lexicographic over the columns: a 32-byte header
 * ("GEMEMB01", u32 num_v, u32 0, u64 rows, u64 digest), then the rows, u32
 * little-endian.  digest = FNV-1a 64 of the row bytes in that order.  Rows are
 * sorted in runs of run_bytes of host memory (0 = 256 MiB); runs beyond the
 * first go to temporary files "<path>.run<i>" and are merged into `path`. */
typedef struct gem_export_result {
  uint64_t rows;    /* embeddings written                        */
  uint64_t digest;  /* FNV-1a 64 of the canonical row bytes       */
  uint32_t runs;    /* sorted runs merged                         */
  uint32_t num_v;
} gem_export_result;
/* Canonical file of n host rows given in any order (host only). */
int gem_export_rows(const uint32_t *rows, uint64_t n, uint32_t num_v, const char *path, uint64_t run_bytes,
                    gem_export_result *out);
/* Every embedding of q (the plain pipeline: filter, tables, GEM order, join)
 * streamed from the GPU into a canonical file.  The join runs without a count
 * cap; more than max_rows embeddings (0 = no bound) stops it with GEM_E_LIMIT
 * and leaves no file, as does the work limit of gem_hip_set_work_limit. */
int gem_hip_export_embeddings(gem_ctx *ctx, const gem_query_view *q, const char *path, uint64_t max_rows,
                              uint64_t run_bytes, gem_export_result *out);

/* Whole plain pipeline for one query: filter -> index -> tables -> GEM order
 * -> join.  Equivalent to the reference's main() for one query at -t 0. */
int gem_hip_match(gem_ctx *ctx, const gem_query_view *q, uint64_t cap,
                  gem_join_result *out, uint32_t *order_out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* GEM_HIP_H */
