// Edge-at-a-time join along a connected edge order (the reference's join():
// in-core/lib/join/src/join_trie.cu:1087-1277 with process() 1025-1085,
// host_select 116-395, direct_join 397-718, refer_join 720-1023).
//
// Rows are materialised level by level as row-major tuples of positions (the
// position of the column's vertex in its query vertex's candidate list), so a
// row's key is read directly instead of climbing the reference's parent-pointer
// chain (join_trie_kernel.cu:124-134) and probing a hash table.
//
//   expand (edge x -> y, y new): row r spawns r + [w] for w in T_{x->y}[r[x]]
//     unless w is the vertex of an earlier column with y's label
//     (injectivity, join_trie.cu:450-462);
//   select (both endpoints present): keep r iff r[m] is in T_{q(k)->q(m)}[r[k]]
//     (sorted list: binary search; join_trie_kernel.cu:148);
//   the last step only counts (join_trie.cu:1204 skip_final_materialization);
//     an expansion's count is |list| minus the equal-label columns found in the
//     list by binary search, so final lists are never walked.
//
// Count pass (one thread per row, 256-row blocks) -> block scan -> write pass
// (waves walk their 64 rows' lists 64 items per step, ballot compaction).  A
// level whose output would exceed the level budget is split into block ranges
// and each range is finished depth-first, so memory stays bounded.
#include <algorithm>
#include <cstring>
#include <functional>

#include "ctx.h"
#include "wave.h"

namespace gemd {

constexpr uint32_t JB = 256;  // rows per block
constexpr uint32_t JW = JB / 64;

struct StepArgs {
  const uint32_t *rows;
  uint64_t nrows;
  uint32_t k;        // input row width
  uint32_t key_col;  // column whose position keys the table
  uint32_t val_col;  // select: tested column
  const uint64_t *toff;
  const uint32_t *tval;
  // expand: injectivity against equal-label columns
  uint32_t ndup;
  uint32_t dup_col[GEM_MAX_QV];
  const uint32_t *dup_can[GEM_MAX_QV];  // Can(q(col)) lists (vertex ids)
  const uint64_t *bm_new;               // bitmap / rank of the new vertex's candidates
  const uint32_t *rank_new;
  uint32_t W;
};

__device__ __forceinline__ uint32_t position_in(const uint64_t *bm, const uint32_t *rank, uint32_t v) {
  const uint64_t word = bm[v >> 6];
  const uint32_t b = v & 63;
  if (!((word >> b) & 1ull)) return NOT_FOUND;
  return rank[v >> 6] + (uint32_t)__popcll(word & ((1ull << b) - 1));
}

// Bytes charged for one binary search over a list of len entries:
// 4 * ceil(log2(len) + 1).
__device__ __forceinline__ uint64_t search_bytes(uint64_t len) {
  if (len == 0) return 0;
  const uint32_t cl = len == 1 ? 0u : 64u - (uint32_t)__clzll(len - 1);
  return 4ull * (cl + 1);
}

// Count of row r's expansion: list length minus equal-label columns that occur
// in the list.  Also returns the list start/length, (optionally) the dup
// positions, and the search bytes spent.
__device__ __forceinline__ uint64_t expand_count_row(const StepArgs &a, uint64_t r, uint64_t *o0, uint64_t *len,
                                                     uint32_t *dup_pos, uint64_t *sbytes) {
  const uint32_t *row = a.rows + r * a.k;
  const uint32_t p = row[a.key_col];
  const uint64_t b = a.toff[p], e = a.toff[p + 1];
  *o0 = b;
  *len = e - b;
  uint64_t cnt = e - b;
  for (uint32_t c = 0; c < a.ndup; ++c) {
    const uint32_t vid = a.dup_can[c][row[a.dup_col[c]]];
    const uint32_t pos = position_in(a.bm_new, a.rank_new, vid);
    if (dup_pos) dup_pos[c] = pos;
    if (pos != NOT_FOUND) {
      *sbytes += search_bytes(e - b);
      if (sorted_contains(a.tval + b, e - b, pos)) --cnt;
    }
  }
  return cnt;
}

__device__ __forceinline__ bool select_row(const StepArgs &a, uint64_t r, uint64_t *len) {
  const uint32_t *row = a.rows + r * a.k;
  const uint32_t p = row[a.key_col], v = row[a.val_col];
  const uint64_t b = a.toff[p], e = a.toff[p + 1];
  *len = e - b;
  return sorted_contains(a.tval + b, e - b, v);
}

__device__ __forceinline__ uint64_t block_sum_u64(uint64_t x, uint64_t *sm) {
  x = wave_sum_u64(x);
  const uint32_t w = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) sm[w] = x;
  __syncthreads();
  uint64_t t = 0;
  for (uint32_t i = 0; i < JW; ++i) t += sm[i];
  __syncthreads();
  return t;
}

// Per-block counts (or, with FINAL, one atomic add of the block sum into
// aux[0]).  With `prof` set, aux[1] += list entries addressed and aux[2] +=
// binary-search bytes (algorithmic-byte bookkeeping).
template <bool SELECT, bool FINAL>
__global__ void __launch_bounds__(JB) k_step_count(const StepArgs a, uint64_t *__restrict__ block_counts,
                                                  unsigned long long *__restrict__ aux, int prof) {
  __shared__ uint64_t sm[JW];
  const uint32_t b = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint64_t r = (uint64_t)b * JB + threadIdx.x;
  uint64_t cnt = 0, len = 0, sb = 0;
  if (r < a.nrows) {
    if (SELECT) {
      cnt = select_row(a, r, &len) ? 1 : 0;
      sb = search_bytes(len);
    } else {
      uint64_t o0;
      cnt = expand_count_row(a, r, &o0, &len, nullptr, &sb);
    }
  }
  const uint64_t s = block_sum_u64(cnt, sm);
  if (threadIdx.x == 0) {
    if (FINAL) {
      if (s) atomicAdd(aux, (unsigned long long)s);
    } else {
      block_counts[b] = s;
    }
  }
  if (prof) {
    const uint64_t ls = block_sum_u64(len, sm), ss = block_sum_u64(sb, sm);
    if (threadIdx.x == 0) {
      atomicAdd(aux + 1, (unsigned long long)ls);
      atomicAdd(aux + 2, (unsigned long long)ss);
    }
  }
}

// Block-level exclusive scan of one u64 per thread.
__device__ __forceinline__ uint64_t block_excl_u64(uint64_t x, uint64_t *sm) {
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint64_t inc = wave_incl_scan_u64(x);
  if (l == 63) sm[w] = inc;
  __syncthreads();
  uint64_t base = 0;
  for (uint32_t i = 0; i < w; ++i) base += sm[i];
  __syncthreads();
  return base + inc - x;
}

// Select write: compact kept rows.  Blocks [b_first, b_first + gridDim.x).
__global__ void __launch_bounds__(JB) k_select_write(const StepArgs a, const uint64_t *__restrict__ block_off,
                                                    uint32_t b_first, uint64_t out_base, uint32_t *__restrict__ out) {
  __shared__ uint64_t sm[JW];
  const uint32_t b = b_first + xcd_swizzle(blockIdx.x, gridDim.x);
  const uint64_t r = (uint64_t)b * JB + threadIdx.x;
  uint64_t len;
  const bool keep = r < a.nrows && select_row(a, r, &len);
  const uint64_t ex = block_excl_u64(keep ? 1 : 0, sm);
  if (keep) {
    const uint64_t o = block_off[b] - out_base + ex;
    const uint32_t *src = a.rows + r * a.k;
    uint32_t *dst = out + o * a.k;
    for (uint32_t c = 0; c < a.k; ++c) dst[c] = src[c];
  }
}

__device__ __forceinline__ uint64_t lanes_between(uint32_t lo, uint32_t hi) {  // bits [lo, hi]
  const uint64_t h = (hi >= 63) ? ~0ull : ((1ull << (hi + 1)) - 1);
  return h & ~((1ull << lo) - 1);
}

// Expand write.  Dynamic LDS: JW * 64 * ndup u32 dup positions.
__global__ void __launch_bounds__(JB) k_expand_write(const StepArgs a, const uint64_t *__restrict__ block_off,
                                                    uint32_t b_first, uint64_t out_base, uint32_t *__restrict__ out,
                                                    unsigned long long *__restrict__ aux, int prof) {
  __shared__ uint64_t sm[JW];
  __shared__ uint32_t st[JW][65];
  __shared__ uint64_t ob[JW][64];
  __shared__ uint64_t ro[JW][64];
  extern __shared__ uint32_t dup_lds[];
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t b = b_first + xcd_swizzle(blockIdx.x, gridDim.x);
  const uint64_t r = (uint64_t)b * JB + threadIdx.x;
  uint32_t *dp = dup_lds + (uint64_t)(w * 64 + l) * a.ndup;
  uint64_t cnt = 0, o0 = 0, len = 0, sb = 0;
  if (r < a.nrows) cnt = expand_count_row(a, r, &o0, &len, dp, &sb);
  const uint64_t ex = block_excl_u64(cnt, sm);
  const uint32_t ilen = (uint32_t)len;
  const uint32_t incl = wave_incl_scan_u32(ilen);
  const uint32_t total = readlane_u32(incl, 63);
  if (prof && l == 0 && total) atomicAdd(aux + 1, (unsigned long long)total);
  st[w][l] = incl - ilen;
  if (l == 63) st[w][64] = total;
  ob[w][l] = o0;
  ro[w][l] = block_off[b] - out_base + ex;
  __syncthreads();
  const uint64_t wave_row0 = (uint64_t)b * JB + w * 64;
  const uint32_t k = a.k;
  uint32_t carry = 0;
  for (uint32_t t = 0; t < total; t += 64) {
    const uint32_t item = t + l;
    const bool iv = item < total;
    const uint32_t row = iv ? seg_of(st[w], item) : 63;
    const uint32_t rs = st[w][row], re = st[w][row + 1];
    const uint32_t last = min(63u, total - 1 - t);
    const uint32_t row0 = readlane_u32(row, 0);
    bool keep = false;
    uint32_t val = 0;
    if (iv) {
      val = a.tval[ob[w][row] + (item - rs)];
      keep = true;
      const uint32_t *rdp = dup_lds + (uint64_t)(w * 64 + row) * a.ndup;
      for (uint32_t c = 0; c < a.ndup; ++c)
        if (rdp[c] == val) keep = false;
    }
    const uint64_t m = __ballot(keep);
    const uint32_t seg_lo = rs > t ? rs - t : 0;
    const uint32_t seg_hi = (re < t + 64 ? re - t : 64) - 1;
    const bool row_done = re <= t + 64;
    const uint32_t before = (row == row0) ? carry : 0;
    if (keep) {
      const uint64_t below = lanes_between(seg_lo, l) & ~(1ull << l);
      const uint64_t o = ro[w][row] + before + (uint64_t)__popcll(m & below);
      const uint32_t *src = a.rows + (wave_row0 + row) * k;
      uint32_t *dst = out + o * (k + 1);
      for (uint32_t c = 0; c < k; ++c) dst[c] = src[c];
      dst[k] = val;
    }
    const uint32_t tot_row = iv ? before + (uint32_t)__popcll(m & lanes_between(seg_lo, seg_hi)) : 0;
    carry = readlane_u32(row_done ? 0u : tot_row, last);
  }
}

// Final rows (positions) -> vertex ids in query-vertex order.
__global__ void k_rows_to_ids(const uint32_t *__restrict__ rows, uint64_t nrows, uint32_t k,
                              const uint32_t *const *__restrict__ col_can, const uint32_t *__restrict__ col_u,
                              uint32_t nv, uint32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows * k) return;
  const uint64_t r = i / k;
  const uint32_t c = (uint32_t)(i % k);
  out[r * nv + col_u[c]] = col_can[c][rows[i]];
}

__global__ void k_iota(uint32_t *p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

struct JStep {
  bool select;
  uint32_t table;  // directed edge
  uint32_t key_col, val_col, new_u;
  uint32_t ndup;
  uint32_t dup_col[GEM_MAX_QV];
};

class JoinRunner {
 public:
  JoinRunner(gem_ctx *c, const gemh::JoinPlan &plan, uint64_t cap, gem_step_log *log, EmbSink *sink)
      : c_(c), Qd_(c->q), cap_(cap), log_(log), sink_(sink) {
    // step 0: the first table seeds 2-column rows from the 1-column identity
    JStep s0{};
    s0.select = false;
    s0.table = plan.first_directed;
    s0.key_col = 0;
    s0.new_u = plan.col_u[1];
    steps_.push_back(s0);
    for (const auto &ps : plan.steps) {
      JStep s{};
      s.select = ps.select;
      s.table = ps.directed;
      s.key_col = ps.key_col;
      s.val_col = ps.val_col;
      s.new_u = ps.new_u;
      s.ndup = ps.num_dup;
      for (uint32_t i = 0; i < ps.num_dup; ++i) s.dup_col[i] = ps.dup_cols[i];
      steps_.push_back(s);
    }
    col_u_ = plan.col_u;
    rows_.assign(steps_.size(), 0);
    ms_.assign(steps_.size(), 0.0f);
    budget_ = std::max<uint64_t>(c->level_budget_bytes, 1ull << 20);
    work_limit_ = c->work_limit_rows;
  }

  void run(JoinOut *out) {
    hipStream_t s = c_->stream;
    d_aux_.alloc(24, s);
    GEM_HIP_CHECK(hipMemsetAsync(d_aux_.p, 0, 24, s));
    GEM_HIP_CHECK(hipEventCreate(&e0_));
    GEM_HIP_CHECK(hipEventCreate(&e1_));
    const uint32_t a = col_u_[0];
    const uint64_t n0 = Qd_.ncan[a];
    DevBuf lvl;
    lvl.alloc(std::max<uint64_t>(n0, 1) * 4, s);
    if (n0) {
      k_iota<<<(uint32_t)((n0 + 255) / 256), 256, 0, s>>>(lvl.as<uint32_t>(), n0);
      GEM_LAUNCH_CHECK("k_iota");
      level(0, lvl.as<uint32_t>(), n0, 1);
    }
    (void)hipEventDestroy(e0_);
    (void)hipEventDestroy(e1_);
    out->count = saturated_ ? cap_ : final_total_;
    out->saturated = saturated_;
    out->emb_written = emb_rows_;
    out->sink_stopped = sink_stopped_;
    out->rows_processed = processed_;
    out->work_limited = work_limited_;
    if (work_limited_) {
      out->count = 0;
      out->saturated = false;
    }
    if (log_) {
      for (size_t i = 0; i < steps_.size() && i < GEM_MAX_QE; ++i) {
        log_->edge[i] = i == 0 ? first_edge_ : steps_[i].table >> 1;
        log_->is_select[i] = steps_[i].select;
        log_->rows[i] = rows_[i];
        log_->ms[i] = ms_[i];
      }
    }
  }

  uint32_t first_edge_ = 0;

 private:
  gem_ctx *c_;
  QueryDev &Qd_;
  uint64_t cap_;
  gem_step_log *log_;
  EmbSink *sink_;
  std::vector<JStep> steps_;
  std::vector<uint32_t> col_u_;
  std::vector<uint64_t> rows_;
  std::vector<float> ms_;
  uint64_t budget_;
  DevBuf d_aux_;
  uint64_t final_total_ = 0;
  bool saturated_ = false;
  bool stop_ = false;
  uint64_t processed_ = 0;
  uint64_t work_limit_ = 0;
  bool work_limited_ = false;
  uint64_t emb_rows_ = 0;
  bool sink_stopped_ = false;
  std::vector<uint32_t> host_rows_;
  hipEvent_t e0_{}, e1_{};

  StepArgs args(const JStep &st, const uint32_t *rows, uint64_t n, uint32_t k) {
    StepArgs a{};
    a.rows = rows;
    a.nrows = n;
    a.k = k;
    a.key_col = st.key_col;
    a.val_col = st.val_col;
    const TableDev &T = Qd_.tab[st.table];
    if (!T.present) throw Error{GEM_E_STATE, "join step uses a table that was not built"};
    a.toff = T.off.as<uint64_t>();
    a.tval = T.val.as<uint32_t>();
    a.ndup = st.ndup;
    a.W = Qd_.W;
    if (!st.select) {
      a.bm_new = Qd_.bitmap.as<uint64_t>() + (uint64_t)st.new_u * Qd_.W;
      a.rank_new = Qd_.rank.as<uint32_t>() + (uint64_t)st.new_u * (Qd_.W + 1);
      for (uint32_t i = 0; i < st.ndup; ++i) {
        a.dup_col[i] = st.dup_col[i];
        a.dup_can[i] = Qd_.can.as<uint32_t>() + Qd_.can_off[col_u_[st.dup_col[i]]];
      }
    }
    return a;
  }

  uint64_t read_u64(const uint64_t *d) {
    uint64_t *h = pinned<uint64_t>(c_, 1);
    GEM_HIP_CHECK(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, c_->stream));
    GEM_HIP_CHECK(hipStreamSynchronize(c_->stream));
    return *h;
  }

  void time_begin() { GEM_HIP_CHECK(hipEventRecord(e0_, c_->stream)); }
  void time_end(uint32_t s) {
    GEM_HIP_CHECK(hipEventRecord(e1_, c_->stream));
    GEM_HIP_CHECK(hipEventSynchronize(e1_));
    float ms = 0;
    GEM_HIP_CHECK(hipEventElapsedTime(&ms, e0_, e1_));
    ms_[s] += ms;
  }

  // aux[0] = final count, aux[1] = list entries addressed, aux[2] = search bytes
  void read_aux(uint64_t *out3) {
    uint64_t *h = pinned<uint64_t>(c_, 3);
    GEM_HIP_CHECK(hipMemcpyAsync(h, d_aux_.p, 24, hipMemcpyDeviceToHost, c_->stream));
    GEM_HIP_CHECK(hipStreamSynchronize(c_->stream));
    out3[0] = h[0];
    out3[1] = h[1];
    out3[2] = h[2];
  }

  void final_count(uint32_t si, const uint32_t *rows, uint64_t n, uint32_t k) {
    const JStep &st = steps_[si];
    StepArgs a = args(st, rows, n, k);
    const uint32_t nb = (uint32_t)((n + JB - 1) / JB);
    const int prof = c_->prof.on ? 1 : 0;
    time_begin();
    GEM_HIP_CHECK(hipMemsetAsync(d_aux_.p, 0, 24, c_->stream));
    {
      ProfLaunch pl(c_, st.select ? K_SELECT_FINAL : K_EXPAND_FINAL);
      if (st.select)
        k_step_count<true, true><<<nb, JB, 0, c_->stream>>>(a, nullptr, d_aux_.as<unsigned long long>(), prof);
      else
        k_step_count<false, true><<<nb, JB, 0, c_->stream>>>(a, nullptr, d_aux_.as<unsigned long long>(), prof);
      GEM_LAUNCH_CHECK("k_step_count<final>");
    }
    uint64_t aux[3];
    read_aux(aux);
    time_end(si);
    // rows + offset pair per row + searches
    prof_bytes(c_, st.select ? K_SELECT_FINAL : K_EXPAND_FINAL, (4.0 * k + 8.0) * n + (double)aux[2]);
    const uint64_t got = aux[0];
    rows_[si] += got;
    final_total_ += got;
    if (cap_ && final_total_ > cap_) {
      saturated_ = true;
      stop_ = true;
    }
  }

  void emit_rows(const uint32_t *rows, uint64_t n, uint32_t k) {
    // rows are final (k == nv columns): vertex ids, in query-vertex order, to the sink
    const uint32_t nv = Qd_.Q.nv;
    if (!sink_ || n == 0) return;
    hipStream_t s = c_->stream;
    std::vector<const uint32_t *> cc(k);
    for (uint32_t c = 0; c < k; ++c) cc[c] = Qd_.can.as<uint32_t>() + Qd_.can_off[col_u_[c]];
    DevBuf dcc, dcu, dout;
    dcc.alloc(k * 8, s);
    dcu.alloc(k * 4, s);
    dout.alloc(n * nv * 4, s);
    GEM_HIP_CHECK(hipMemcpyAsync(dcc.p, cc.data(), k * 8, hipMemcpyHostToDevice, s));
    GEM_HIP_CHECK(hipMemcpyAsync(dcu.p, col_u_.data(), k * 4, hipMemcpyHostToDevice, s));
    const uint64_t tot = n * k;
    k_rows_to_ids<<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(rows, n, k, dcc.as<const uint32_t *>(),
                                                                 dcu.as<uint32_t>(), nv, dout.as<uint32_t>());
    GEM_LAUNCH_CHECK("k_rows_to_ids");
    host_rows_.resize(n * nv);
    GEM_HIP_CHECK(hipMemcpyAsync(host_rows_.data(), dout.p, n * nv * 4, hipMemcpyDeviceToHost, s));
    GEM_HIP_CHECK(hipStreamSynchronize(s));
    emb_rows_ += n;
    if (!sink_->put(host_rows_.data(), n, nv)) {
      sink_stopped_ = true;
      stop_ = true;
    }
  }

  // Apply step si to `n` rows of width k.
  void level(uint32_t si, const uint32_t *rows, uint64_t n, uint32_t k) {
    if (stop_ || n == 0) return;
    processed_ += n;
    if (work_limit_ && processed_ > work_limit_) {
      work_limited_ = true;
      stop_ = true;
      return;
    }
    const JStep &st = steps_[si];
    const bool last = si + 1 == steps_.size();
    if (last && !sink_) {
      final_count(si, rows, n, k);
      return;
    }
    hipStream_t s = c_->stream;
    StepArgs a = args(st, rows, n, k);
    const uint32_t nb = (uint32_t)((n + JB - 1) / JB);
    const int prof = c_->prof.on ? 1 : 0;
    time_begin();
    DevBuf bc, bo;
    bc.alloc((uint64_t)nb * 8, s);
    bo.alloc(((uint64_t)nb + 1) * 8, s);
    if (prof) GEM_HIP_CHECK(hipMemsetAsync(d_aux_.p, 0, 24, s));
    {
      ProfLaunch pl(c_, st.select ? K_SELECT_COUNT : K_EXPAND_COUNT);
      if (st.select)
        k_step_count<true, false><<<nb, JB, 0, s>>>(a, bc.as<uint64_t>(), d_aux_.as<unsigned long long>(), prof);
      else
        k_step_count<false, false><<<nb, JB, 0, s>>>(a, bc.as<uint64_t>(), d_aux_.as<unsigned long long>(), prof);
      GEM_LAUNCH_CHECK("k_step_count");
    }
    scan_u64(c_, bc.as<uint64_t>(), bo.as<uint64_t>(), nb);
    const uint64_t total = read_u64(bo.as<uint64_t>() + nb);
    bc.reset();
    rows_[si] += total;
    const uint32_t ko = st.select ? k : k + 1;
    const uint64_t row_bytes = (uint64_t)ko * 4;
    double search_per_row = 0;
    if (prof) {
      uint64_t aux[3];
      read_aux(aux);
      // count pass: rows, offset pairs, searches
      prof_bytes(c_, st.select ? K_SELECT_COUNT : K_EXPAND_COUNT, (4.0 * k + 8.0) * n + (double)aux[2]);
      search_per_row = n ? (double)aux[2] / (double)n : 0.0;
    }
    if (total == 0) {
      time_end(si);
      return;
    }
    // block ranges whose output fits the level budget
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
    std::vector<uint64_t> hoff;
    if (total * row_bytes <= budget_) {
      ranges.push_back({0, nb});
    } else {
      hoff.resize((size_t)nb + 1);
      GEM_HIP_CHECK(hipMemcpyAsync(hoff.data(), bo.p, ((uint64_t)nb + 1) * 8, hipMemcpyDeviceToHost, s));
      GEM_HIP_CHECK(hipStreamSynchronize(s));
      const uint64_t lim = std::max<uint64_t>(budget_ / row_bytes, 1);
      uint32_t b0 = 0;
      while (b0 < nb) {
        uint32_t b1 = b0 + 1;
        while (b1 < nb && hoff[b1 + 1] - hoff[b0] <= lim) ++b1;
        ranges.push_back({b0, b1});
        b0 = b1;
      }
    }
    bool timed = false;
    for (auto [b0, b1] : ranges) {
      if (stop_) break;
      const uint64_t base = hoff.empty() ? 0 : hoff[b0];
      const uint64_t cnt = hoff.empty() ? total : hoff[b1] - hoff[b0];
      if (cnt == 0) continue;
      if (timed) time_begin();
      DevBuf outb;
      outb.alloc(cnt * row_bytes, s);
      const uint32_t g = b1 - b0;
      if (prof) GEM_HIP_CHECK(hipMemsetAsync(d_aux_.p, 0, 24, s));
      {
        ProfLaunch pl(c_, st.select ? K_SELECT_WRITE : K_EXPAND_WRITE);
        if (st.select) {
          k_select_write<<<g, JB, 0, s>>>(a, bo.as<uint64_t>(), b0, base, outb.as<uint32_t>());
          GEM_LAUNCH_CHECK("k_select_write");
        } else {
          const size_t lds = (size_t)JW * 64 * st.ndup * 4;
          k_expand_write<<<g, JB, lds, s>>>(a, bo.as<uint64_t>(), b0, base, outb.as<uint32_t>(),
                                            d_aux_.as<unsigned long long>(), prof);
          GEM_LAUNCH_CHECK("k_expand_write");
        }
      }
      if (prof) {
        // rows of the range (+ their searches), list entries walked, rows written
        uint64_t aux[3];
        read_aux(aux);
        const double nr = (double)std::min<uint64_t>((uint64_t)(b1 - b0) * JB, n - (uint64_t)b0 * JB);
        prof_bytes(c_, st.select ? K_SELECT_WRITE : K_EXPAND_WRITE,
                   (4.0 * k + 8.0 + search_per_row) * nr + 4.0 * (double)aux[1] + (double)row_bytes * cnt);
      }
      time_end(si);
      timed = true;
      if (last) {
        emit_rows(outb.as<uint32_t>(), cnt, ko);
        final_total_ += cnt;
      } else {
        level(si + 1, outb.as<uint32_t>(), cnt, ko);
      }
    }
    if (!timed) time_end(si);
  }
};

// Keeps the first `cap` rows in a host array; later rows are dropped and
// mark the array overflowed (the join goes on and counts them).
struct ArraySink : EmbSink {
  uint32_t *out;
  uint64_t cap, rows = 0;
  bool overflow = false;
  ArraySink(uint32_t *o, uint64_t c) : out(o), cap(c) {}
  bool put(const uint32_t *r, uint64_t n, uint32_t nv) override {
    if (overflow || rows + n > cap) {
      overflow = true;
      return true;
    }
    std::memcpy(out + rows * nv, r, n * nv * 4);
    rows += n;
    return true;
  }
};

void run_join_sink(gem_ctx *c, const uint32_t *order, uint64_t cap, JoinOut *out, gem_step_log *log,
                   EmbSink *sink) {
  QueryDev &Qd = c->q;
  if (!Qd.tabled) throw Error{GEM_E_STATE, "edge tables have not been built"};
  gemh::JoinPlan plan;
  if (gemh::make_join_plan(Qd.Q, order, &plan)) throw Error{GEM_E_INVALID, "edge order is not a connected order"};
  hipStream_t s = c->stream;
  GEM_HIP_CHECK(hipEventRecord(c->ev[4], s));
  JoinRunner jr(c, plan, sink ? 0 : cap, log, sink);
  jr.first_edge_ = plan.first_edge;
  jr.run(out);
  GEM_HIP_CHECK(hipEventRecord(c->ev[5], s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  GEM_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
  c->times.join_ms = ms;
  prof_flush(c);
}

void run_join(gem_ctx *c, const uint32_t *order, uint64_t cap, JoinOut *out, gem_step_log *log, uint32_t *emb_out,
              uint64_t emb_capacity) {
  if (!emb_out) {
    run_join_sink(c, order, cap, out, log, nullptr);
    return;
  }
  ArraySink sink(emb_out, emb_capacity);
  run_join_sink(c, order, cap, out, log, &sink);
  out->emb_written = sink.overflow ? 0 : sink.rows;
}

}  // namespace gemd
