// Device context, per-query device state and shared helpers of the HIP
// library.  gfx950 only (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "../host/planner.h"
#include "../host/pool_core.h"
#include "../host/vsched.h"
#include "gem_hip.h"

namespace gemd {

constexpr uint32_t WAVE = 64;
constexpr uint32_t NO_SLOT = 0xFFu;
constexpr uint32_t NOT_FOUND = 0xFFFFFFFFu;

// Thrown inside the library, turned into a status + message at the C ABI.
struct Error {
  int code;
  std::string msg;
};

#define GEM_HIP_CHECK(call)                                                              \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::gemd::Error{GEM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// Launch-error check after a kernel launch.
#define GEM_LAUNCH_CHECK(name)                                                            \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess)                                                                 \
      throw ::gemd::Error{GEM_E_HIP, std::string("launch ") + (name) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// Device memory of DevBuf comes from a caching allocator over hipMalloc
// (host/pool_core.h): small requests from size classes, large ones carved
// from 8 GiB segments; freed memory is handed out again to allocations on the
// stream that freed it.  All device work of a context runs on its one stream,
// so a block the host frees after enqueueing its last use can serve later
// work on that stream without a wait.  A failed hipMalloc releases every
// cached block and empty segment and retries.
void *pool_alloc(size_t n, hipStream_t s, size_t *cls);
void pool_free(void *p, size_t cls, hipStream_t s);
// hipFree every cached block of stream s (its work must have finished).
void pool_release(hipStream_t s);
// Bytes the allocator holds that no buffer uses and that a large request can
// use whole: empty segments and parked small blocks.
size_t pool_reusable_bytes();
// Return empty segments and parked small blocks to the driver (waits for the
// device); the bytes released.
uint64_t pool_trim(uint64_t retain = 0);
// In-use / peak / held bytes of the allocator (optionally restarting the peak).
gemh::PoolStats pool_stats(bool reset_peak);

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  size_t cls = 0;  // size of the pool block behind p
  hipStream_t s = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), cls(o.cls), s(o.s) { o.p = nullptr; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      cls = o.cls;
      s = o.s;
      o.p = nullptr;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  void alloc(size_t n, hipStream_t stream) {
    reset();
    s = stream;
    if (n == 0) return;
    p = pool_alloc(n, stream, &cls);
    bytes = n;
  }
  void reset() {
    if (p) pool_free(p, cls, s);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// Data graph resident in HBM.  Labels are replaced by dense label ids
// (first-seen order over the vertex array) held as u8 when the graph has at
// most 255 distinct labels, else u32.
// Pinned host allocation (hipHostMalloc), move-only.
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  PinnedBuf(PinnedBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; }
  PinnedBuf &operator=(PinnedBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
    }
    return *this;
  }
  ~PinnedBuf() { reset(); }
  void alloc(size_t n, unsigned flags = hipHostMallocDefault) {
    reset();
    if (n == 0) return;
    hipError_t e = hipHostMalloc(&p, n, flags);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{GEM_E_NOMEM, "hipHostMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e)};
    }
    bytes = n;
  }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

struct GraphDev {
  bool loaded = false;
  uint32_t n = 0;
  uint64_t slots = 0;
  uint64_t max_deg = 0;
  uint32_t n_hub = 0;  // vertices of degree > FILTER_HUB_DEG
  DevBuf off;     // u64[n + 1]
  DevBuf nbr;     // u32[slots] in HBM, or
  PinnedBuf nbr_host;  // u32[slots] in mapped pinned host memory (host-resident graph)
  const uint32_t *nbr_mapped = nullptr;  // device address of nbr_host
  DevBuf lab;     // u8[n] or u32[n] dense label ids
  bool lab8 = true;
  std::unordered_map<uint32_t, uint32_t> label_id;  // original -> dense
  std::vector<uint64_t> label_freq;                 // dense id -> #vertices
  std::vector<uint64_t> label_deg;                  // dense id -> sum of degrees
  // Neighbour array as the kernels address it: HBM, or host memory read over
  // PCIe by the filter and table kernels.
  const uint32_t *nbr_dev() const { return nbr_mapped ? nbr_mapped : nbr.as<uint32_t>(); }
};

// The filter walks the adjacency of vertices of larger degree than this in
// chunks of this many entries spread over waves.
constexpr uint32_t FILTER_HUB_DEG = 4096;

// Query description consumed by the filter kernel (device copy).
struct FilterQuery {
  uint32_t nv, nslots, T, W;
  uint32_t qdeg[GEM_MAX_QV];
  uint32_t group[GEM_MAX_QV];   // slot -> mask of query vertices with that label
  uint32_t nreq[GEM_MAX_QV];
  uint8_t req_slot[GEM_MAX_QV][GEM_MAX_QV];
  uint8_t req_cnt[GEM_MAX_QV][GEM_MAX_QV];
};

struct TableDev {
  bool present = false;
  uint32_t key_u = 0, val_u = 0;
  uint64_t size = 0;
  DevBuf off;  // u64[|Can(key)| + 1]
  DevBuf val;  // u32[size], positions in Can(val_u); empty once streamed
  // streamed form (values in pinned host memory): keys [b << kshift,
  // (b + 1) << kshift) form block b, whose values are blk_off[b]..blk_off[b+1]
  const uint32_t *hval = nullptr;
  uint32_t kshift = 0;
  std::vector<uint64_t> blk_off;  // nblk + 1 entries
};

struct QueryDev {
  bool filtered = false, tabled = false;
  gemh::QueryInfo Q;
  uint32_t W = 0;                      // 64-bit words per candidate bitmap
  std::vector<uint32_t> ncan;          // |Can(u)|
  std::vector<uint64_t> can_off;       // start of Can(u) inside `can`
  DevBuf can;                          // u32 candidate lists (sorted)
  DevBuf bitmap;                       // u64[nv * W]
  DevBuf rank;                         // u32[nv * (W + 1)]: exclusive popcount prefix
  DevBuf fq;                           // FilterQuery
  DevBuf slot_of;                      // u8[num data labels]: dense label -> query slot
  std::vector<TableDev> tab;           // per directed edge
  bool streamed = false;               // table values live in host_vals
  PinnedBuf host_vals;
};

// Kernel-level profiling: HIP events around launches on the context stream
// plus the algorithmic bytes each launch must move (compulsory traffic: every
// input element read once, every output element written once).
enum KernelId {
  K_FILTER,
  K_FILTER_HEAVY,
  K_COMPACT,
  K_TABLES_COUNT,
  K_TABLES_WRITE,
  K_EXPAND_COUNT,
  K_SELECT_COUNT,
  K_EXPAND_FINAL,
  K_SELECT_FINAL,
  K_EXPAND_WRITE,
  K_SELECT_WRITE,
  K_VLEN,
  K_VEXT_WRITE,
  K_VEXT_FINAL,
  K_VCOUNT1,
  K_VEXT_HEAVY_WRITE,
  K_VEXT_HEAVY_FINAL,
  K_NUM
};
inline const char *kernel_name(int id) {
  static const char *n[K_NUM] = {"k_filter", "k_filter_heavy", "k_compact", "k_tables<count>",
                                 "k_tables<write>",    "k_step_count<expand>", "k_step_count<select>",
                                 "k_step_count<expand,final>", "k_step_count<select,final>", "k_expand_write",
                                 "k_select_write", "k_vlen", "k_vext<write>", "k_vext<final>", "k_vcount1",
                                 "k_vext_heavy<write>", "k_vext_heavy<final>"};
  return n[id];
}
struct KStat {
  uint64_t launches = 0;
  double ms = 0, bytes = 0;
};
struct Prof {
  bool on = false;
  KStat st[K_NUM];
  struct Pending {
    int id;
    hipEvent_t b, e;
  };
  std::vector<Pending> pend;
  std::vector<hipEvent_t> pool;
};

}  // namespace gemd

struct gem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  std::string err;
  gemd::GraphDev g;
  gemd::QueryDev q;
  gem_phase_times times{};
  uint64_t level_budget_bytes = 1ull << 31;  // per join level buffer
  uint64_t work_limit_rows = 0;             // 0 = none
  uint32_t persistent_blocks = 2048;        // grid of the batched path's persistent kernels
  uint32_t heavy_t = 1024, heavy_chunk = 1024;  // batched walks: long driving lists split over waves
  uint32_t max_active = 64;                 // batched path: queries running at once (0 = all)
  int graph_host = 0;                       // next data graph: neighbours stay in pinned host memory
  bool relaxed_share = true;                // relaxed queries share one filter and one table build
  bool racing = true;                       // batched learned path: twins race stragglers
  bool prep_overlap = true;                 // batched learned path: later queries prepared during the join
  // A context that prepares queries on its own stream for another one reads
  // that context's data graph (gsrc) instead of its own; `prep` is that
  // context of this one (made at its first use, destroyed with this one, so
  // that its stream's pool memory is reused from run to run).
  const gemd::GraphDev *gsrc = nullptr;
  gem_ctx *prep = nullptr;
  uint64_t swap_bytes = 0;                  // batched paths: host memory pending levels may move to
  bool use_graph = false;                   // replay the batched extend step as a hipGraph
  // batched path with streamed edge tables: values in pinned host memory,
  // a device cache of stream_cache_bytes in slots of stream_slot_bytes
  // (0 = tables stay in HBM)
  uint64_t stream_cache_bytes = 0;
  uint64_t stream_slot_bytes = 1ull << 20;
  gem_stream_stats sstats{};
  void *pinned = nullptr;                   // small pinned staging area
  size_t pinned_bytes = 0;
  gemd::Prof prof;
  // the last batched run's level budget and scheduler statistics
  uint64_t last_budget = 0;
  gemh::SchedStats last_sched{};
  // order models created on this context: destroying the context releases
  // their device memory and detaches them (a later gem_mlp_destroy only frees
  // the host object, every other gem_mlp_* call returns GEM_E_STATE)
  std::vector<gem_mlp *> mlps;
};

namespace gemd {
void mlp_detach(gem_mlp *m);  // frees m's device buffers, sets m->ctx = nullptr
float mlp_margin(const gem_mlp *m);
}

namespace gemd {

// Event pair around one launch; bytes are added separately (they may only be
// known after a later count).
inline hipEvent_t prof_event(gem_ctx *c) {
  if (!c->prof.pool.empty()) {
    hipEvent_t e = c->prof.pool.back();
    c->prof.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  GEM_HIP_CHECK(hipEventCreate(&e));
  return e;
}
struct ProfLaunch {
  gem_ctx *c;
  int id;
  hipEvent_t b = nullptr;
  ProfLaunch(gem_ctx *ctx, int kid) : c(ctx), id(kid) {
    if (c->prof.on && kid >= 0) {
      b = prof_event(c);
      GEM_HIP_CHECK(hipEventRecord(b, c->stream));
    }
  }
  ~ProfLaunch() {
    if (b) {
      hipEvent_t e = nullptr;
      try {
        e = prof_event(c);
      } catch (...) {
        return;
      }
      (void)hipEventRecord(e, c->stream);
      c->prof.pend.push_back({id, b, e});
    }
  }
};
inline void prof_bytes(gem_ctx *c, int id, double bytes) {
  if (c->prof.on) c->prof.st[id].bytes += bytes;
}
// Resolve finished event pairs (call after a stream synchronisation).
inline void prof_flush(gem_ctx *c) {
  for (auto &p : c->prof.pend) {
    float ms = 0;
    if (hipEventSynchronize(p.e) == hipSuccess && hipEventElapsedTime(&ms, p.b, p.e) == hipSuccess) {
      c->prof.st[p.id].ms += ms;
      c->prof.st[p.id].launches += 1;
    }
    c->prof.pool.push_back(p.b);
    c->prof.pool.push_back(p.e);
  }
  c->prof.pend.clear();
}

// Pinned staging for small device->host reads.
template <class T>
T *pinned(gem_ctx *c, size_t n) {
  size_t need = n * sizeof(T);
  if (need > c->pinned_bytes) {
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    size_t sz = std::max<size_t>(need, 1 << 20);
    GEM_HIP_CHECK(hipHostMalloc(&c->pinned, sz, hipHostMallocDefault));
    c->pinned_bytes = sz;
  }
  return reinterpret_cast<T *>(c->pinned);
}

// ---------------------------------------------------------------- phases
// The data graph a context's filter and table kernels read.
inline const GraphDev &graph(const gem_ctx *c) { return c->gsrc ? *c->gsrc : c->g; }

void run_filter(gem_ctx *c, const gem_query_view *qv, uint32_t T);
void run_tables(gem_ctx *c);
// the same phases on an explicit per-query state (several queries resident)
void run_filter(gem_ctx *c, QueryDev &Qd, const gem_query_view *qv, uint32_t T);
// The filter into caller-owned device arrays with the library's fixed scratch
// array (no allocation); filter_scratch_init finds that array's address on
// the current device (once per device, at context creation).
void filter_scratch_init(int device);
uint64_t filter_scratch_bytes(const GraphDev &G, uint32_t nv);
void run_filter_noalloc(gem_ctx *c, const gem_query_view *qv, uint32_t T, uint32_t *d_num_can, uint32_t *d_can,
                        uint64_t can_pitch);
void run_tables(gem_ctx *c, QueryDev &Qd);
void run_position_lookup(gem_ctx *c, uint32_t u, const uint32_t *h_vertices, uint64_t n, uint32_t *h_pos);

struct JoinOut {
  uint64_t count = 0;
  bool saturated = false;
  uint64_t emb_written = 0;
  uint64_t rows_processed = 0;
  bool work_limited = false;
  bool sink_stopped = false;  // the sink refused rows and the join stopped
};

// Receives a join's final rows as vertex ids (n rows of nv ids, column u =
// query vertex u) in host memory; returning false stops the join.
struct EmbSink {
  virtual ~EmbSink() = default;
  virtual bool put(const uint32_t *rows, uint64_t n, uint32_t nv) = 0;
};
// The plain join with its final rows going to `sink` (count cap off); null
// sink: count only, capped at cap.
void run_join_sink(gem_ctx *c, const uint32_t *order, uint64_t cap, JoinOut *out, gem_step_log *log,
                   EmbSink *sink);
void run_join(gem_ctx *c, const uint32_t *order, uint64_t cap, JoinOut *out, gem_step_log *log, uint32_t *emb_out,
              uint64_t emb_capacity);

void run_match_batch(gem_ctx *c, const gem_query_view *qs, uint32_t nq, const uint32_t *const *vorders,
                     const gem_batch_opts *opts, gem_join_result *results, uint64_t *level_rows);
void run_vjoin(gem_ctx *c, const uint32_t *vorder, uint64_t cap, gem_join_result *out, uint64_t *level_rows);
void run_order_candidates(gem_ctx *c, uint32_t max_orders, uint64_t seed, uint32_t *orders, int32_t *feats,
                          uint32_t *n);
void run_match_batch_learned(gem_ctx *c, const gem_query_view *qs, uint32_t nq, gem_mlp *mlp, uint32_t max_orders,
                             const gem_batch_opts *opts, gem_join_result *results, uint64_t *level_rows,
                             uint32_t *chosen);
void run_match_relaxed(gem_ctx *c, const gem_query_view *q, uint32_t threshold,
                       const std::vector<gemh::EdgeSet> &sets, gem_mlp *mlp, uint32_t max_orders,
                       const gem_batch_opts *opts, gem_join_result *results);

// ------------------------------------------------------------------ scans
// Exclusive scan of n u64 values in place-free form: out[i] = sum(in[0..i)),
// out[n] = total.  Returns nothing; caller reads out[n] if needed.
void scan_u64(gem_ctx *c, const uint64_t *d_in, uint64_t *d_out, uint64_t n);
// Independent in-place exclusive scans of u64 arrays p[0..n) (the total goes
// to p[n] and to d_totals[i]); d_segs is the device copy of h_segs.
struct ScanSeg {
  uint64_t *p;
  uint64_t n;
};
void scan_many_u64(gem_ctx *c, const ScanSeg *d_segs, const ScanSeg *h_segs, uint32_t nseg, uint64_t *d_totals);
// Segmented exclusive scan of u32 counts: nseg segments of length len each,
// out has len + 1 entries per segment (last = segment total).
void scan_segments_u32(gem_ctx *c, const uint32_t *d_in, uint32_t *d_out, uint32_t nseg, uint64_t len);

}  // namespace gemd
