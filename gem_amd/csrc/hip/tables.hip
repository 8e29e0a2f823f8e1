// Edge-candidate tables (the reference's construct_edge_candidates,
// in-core/lib/structure/src/unifiedTrie.cu:1020-1391 with kernels
// construct_kernel.cu:1183-1340; CPU form out-of-core/.../unifiedTrie.cu:1224-1427).
//
// For every query vertex u that is not a leaf and every query neighbour u_j of
// u, table T_{u->u_j} has one row per v in Can(u) (row i <-> Can(u)[i]) holding
// the neighbours w of v with w in Can(u_j), in adjacency order.  Values are
// stored as positions of w in Can(u_j) (pos = rank in u_j's candidate bitmap),
// so join rows carry positions and never need a hash probe.
//
// One pass over adjacency serves all of u's tables: a wave takes 64
// consecutive candidates of one u, flattens their adjacency lists into one
// item range (wave scan of degrees) and walks it 64 items per step; item ->
// row by binary lifting over the 64 row starts in LDS; membership of w in each
// Can(u_j) is one bit of u_j's bitmap (L2-resident); per-row counts and
// in-row ranks come from ballots restricted to the row's lane segment, with
// one carried count for the row that spans two steps.  Count pass -> per-table
// exclusive scan -> write pass.
#include "ctx.h"
#include "wave.h"

namespace gemd {

constexpr uint32_t TAB_WAVES = 4;

struct TableJob {
  uint32_t nu;
  uint32_t W;
  uint32_t wave_start[GEM_MAX_QV + 1];
  uint32_t ncan[GEM_MAX_QV];
  uint64_t can_off[GEM_MAX_QV];
  uint32_t nd[GEM_MAX_QV];    // tables keyed on this u
  uint32_t tab0[GEM_MAX_QV];  // first entry in the TableDesc list
};

struct TableDesc {
  uint32_t val_u;
  uint64_t *off;
  uint32_t *val;
};

__device__ __forceinline__ uint64_t lane_range(uint32_t a, uint32_t b) {  // bits [a, b], a <= b < 64
  const uint64_t hi = (b >= 63) ? ~0ull : ((1ull << (b + 1)) - 1);
  return hi & ~((1ull << a) - 1);
}

// Rows longer than HEAVY_ROW are walked by k_tables_heavy in fixed-size
// chunks spread over waves, so that neither a wave of 64 hub rows nor one
// hub's row can become a straggler.
constexpr uint32_t HEAVY_ROW = 256;

template <bool WRITE>
__global__ void __launch_bounds__(TAB_WAVES * 64)
k_tables(const uint64_t *__restrict__ goff, const uint32_t *__restrict__ gnbr, const uint32_t *__restrict__ can,
         const uint64_t *__restrict__ bitmap, const uint32_t *__restrict__ rank, const TableJob job,
         const TableDesc *__restrict__ desc, unsigned long long *__restrict__ adj_len,
         uint64_t *__restrict__ heavy_list, uint32_t *__restrict__ heavy_n) {
  __shared__ uint32_t st[TAB_WAVES][65];
  __shared__ uint64_t ob[TAB_WAVES][64];
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t gw = blockIdx.x * TAB_WAVES + w;
  if (gw >= job.wave_start[job.nu]) return;
  uint32_t ju = 0;
  while (job.wave_start[ju + 1] <= gw) ++ju;
  const uint32_t row_base = (gw - job.wave_start[ju]) * 64;
  const uint32_t ncan = job.ncan[ju];
  const uint32_t i = row_base + l;
  const bool valid = i < ncan;
  uint64_t o0 = 0;
  uint32_t len = 0, full_len = 0;
  if (valid) {
    const uint32_t v = can[job.can_off[ju] + i];
    o0 = goff[v];
    full_len = (uint32_t)(goff[v + 1] - o0);
    len = full_len > HEAVY_ROW ? 0u : full_len;
  }
  if (!WRITE) {
    const uint64_t hv = __ballot(full_len > HEAVY_ROW);
    if (hv) {
      uint32_t base = 0;
      if (l == 0) base = atomicAdd(heavy_n, (uint32_t)__popcll(hv));
      base = readlane_u32(base, 0);
      if ((hv >> l) & 1ull) heavy_list[base + rank_below(hv)] = ((uint64_t)ju << 32) | i;
    }
  }
  const uint32_t incl = wave_incl_scan_u32(len);
  const uint32_t total = readlane_u32(incl, 63);
  if (!WRITE && l == 0 && total) atomicAdd(adj_len + ju, (unsigned long long)total);
  st[w][l] = incl - len;
  ob[w][l] = o0;
  if (l == 63) st[w][64] = total;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint32_t nd = job.nd[ju], t0 = job.tab0[ju];
  if (!WRITE && valid && full_len == 0)
    for (uint32_t j = 0; j < nd; ++j) desc[t0 + j].off[i] = 0;
  const uint32_t W = job.W;
  // lane j < nd holds table j: value vertex, row offsets, values
  uint32_t my_vu = 0;
  uint64_t my_off = 0, my_val = 0;
  if (l < nd) {
    const TableDesc d = desc[t0 + l];
    my_vu = d.val_u;
    my_off = reinterpret_cast<uint64_t>(d.off);
    my_val = reinterpret_cast<uint64_t>(d.val);
  }
  uint32_t carry_reg = 0;  // lane j: hits of the row carried into the next group, table j
  // One 64-item group of the flattened adjacency: the lane's row, its value,
  // and the lane range of its row inside the group.
  struct Grp {
    bool iv, row_done;
    uint32_t row, wv, last, row0, seg_hi;
    uint64_t seg_mask, below_mask;
  };
  auto group = [&](uint32_t t) {
    Grp g;
    const uint32_t item = t + l;
    g.iv = item < total;
    g.row = g.iv ? seg_of(st[w], item) : 63;
    const uint32_t rs = st[w][g.row], re = st[w][g.row + 1];
    g.wv = g.iv ? gnbr[ob[w][g.row] + (item - rs)] : 0;
    g.last = min(63u, total - 1 - t);
    g.row0 = readlane_u32(g.row, 0);
    const uint32_t seg_lo = rs > t ? rs - t : 0;
    g.seg_hi = (re < t + 64 ? re - t : 64) - 1;
    g.row_done = re <= t + 64;
    g.seg_mask = g.iv ? lane_range(seg_lo, g.seg_hi) : 0;
    g.below_mask = g.iv ? (lane_range(seg_lo, l) & ~(1ull << l)) : 0;
    return g;
  };
  // table j over one group, `carry` = the row's hits before the group; returns
  // the hits carried into the next group
  auto apply = [&](const Grp &g, uint32_t j, uint32_t vu, uint64_t word, uint32_t carry) {
    const bool hit = (word >> (g.wv & 63)) & 1ull;
    const uint64_t m = __ballot(hit);
    const uint32_t before = (g.row == g.row0) ? carry : 0;
    const uint32_t tot_row = before + (uint32_t)__popcll(m & g.seg_mask);
    uint64_t *off = reinterpret_cast<uint64_t *>(readlane_u64(my_off, j));
    if (WRITE) {
      if (hit) {
        uint32_t *val = reinterpret_cast<uint32_t *>(readlane_u64(my_val, j));
        const uint32_t r = before + (uint32_t)__popcll(m & g.below_mask);
        const uint32_t p = rank[(uint64_t)vu * (W + 1) + (g.wv >> 6)] +
                           (uint32_t)__popcll(word & ((1ull << (g.wv & 63)) - 1));
        val[off[row_base + g.row] + r] = p;
      }
    } else {
      if (g.iv && l == g.seg_hi && g.row_done) off[row_base + g.row] = tot_row;
    }
    return readlane_u32(g.row_done ? 0u : tot_row, g.last);
  };
  // two groups per step: both groups' loads are in flight before either
  // group's ballots
  for (uint32_t t = 0; t < total; t += 128) {
    const bool two = t + 64 < total;
    const Grp A = group(t);
    const Grp B = two ? group(t + 64) : A;
    for (uint32_t j = 0; j < nd; ++j) {
      const uint32_t vu = readlane_u32(my_vu, j);
      const uint64_t *bm = bitmap + (uint64_t)vu * W;
      const uint64_t wa = A.iv ? bm[A.wv >> 6] : 0;
      const uint64_t wb = two && B.iv ? bm[B.wv >> 6] : 0;
      uint32_t carry = apply(A, j, vu, wa, readlane_u32(carry_reg, j));
      if (two) carry = apply(B, j, vu, wb, carry);
      if (l == j) carry_reg = carry;
    }
  }
}

// Heavy rows are walked in chunks of TCH adjacency entries, one wave task per
// chunk, so that a hub's row is spread over many waves.  The count pass stores
// each chunk's hit count per table (ccnt[chunk * ndmax + j]); k_theavy_bases
// turns those into each chunk's offset inside its row (exclusive over the
// row's chunks) and writes the rows' totals; the write pass starts each
// chunk's hits at that offset, so values stay in adjacency order.
constexpr uint32_t TCH = 2048;

// Chunk prefix of the heavy rows: cpre[h] = chunks of rows before h,
// cpre[n] = *total = all chunks (one block).
__global__ void __launch_bounds__(1024) k_theavy_prefix(const uint64_t *__restrict__ goff,
                                                       const uint32_t *__restrict__ can, const TableJob job,
                                                       const uint64_t *__restrict__ heavy_list,
                                                       const uint32_t *__restrict__ heavy_n, uint64_t *__restrict__ cpre,
                                                       uint64_t *__restrict__ total) {
  __shared__ uint64_t sm[16];
  const uint32_t n = *heavy_n;
  uint64_t carry = 0;
  for (uint32_t b = 0; b < n; b += blockDim.x) {
    const uint32_t h = b + threadIdx.x;
    uint64_t x = 0;
    if (h < n) {
      const uint64_t e = heavy_list[h];
      const uint32_t v = can[job.can_off[(uint32_t)(e >> 32)] + (uint32_t)e];
      x = (goff[v + 1] - goff[v] + TCH - 1) / TCH;
    }
    uint64_t ex;
    const uint64_t tot = block_excl_scan_u64(x, &ex, sm);
    if (h < n) cpre[h] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    cpre[n] = carry;
    *total = carry;
  }
}

// Persistent grid over the chunks; lane j carries table j's running count.
template <bool WRITE>
__global__ void __launch_bounds__(TAB_WAVES * 64)
k_tables_heavy(const uint64_t *__restrict__ goff, const uint32_t *__restrict__ gnbr, const uint32_t *__restrict__ can,
               const uint64_t *__restrict__ bitmap, const uint32_t *__restrict__ rank, const TableJob job,
               const TableDesc *__restrict__ desc, unsigned long long *__restrict__ adj_len,
               const uint64_t *__restrict__ heavy_list, const uint32_t *__restrict__ heavy_n,
               const uint64_t *__restrict__ cpre, uint32_t *__restrict__ ccnt, uint32_t ndmax) {
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t nh = *heavy_n, W = job.W;
  const uint64_t nchunks = cpre[nh];
  for (uint64_t c = (uint64_t)blockIdx.x * TAB_WAVES + w; c < nchunks; c += (uint64_t)gridDim.x * TAB_WAVES) {
    uint32_t lo = 0, hi = nh;  // the chunk's row: largest h with cpre[h] <= c
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (cpre[mid] <= c)
        lo = mid;
      else
        hi = mid;
    }
    const uint64_t e = heavy_list[lo];
    const uint32_t ju = (uint32_t)(e >> 32), i = (uint32_t)e;
    const uint32_t v = can[job.can_off[ju] + i];
    const uint64_t o0 = goff[v], len = goff[v + 1] - o0;
    const uint64_t t0 = (c - cpre[lo]) * TCH, t1 = t0 + TCH < len ? t0 + TCH : len;
    if (!WRITE && l == 0) atomicAdd(adj_len + ju, (unsigned long long)(t1 - t0));
    const uint32_t nd = job.nd[ju], tb = job.tab0[ju];
    // lane j < nd holds table j: its value vertex, (write) the address of the
    // row's first value, and its running count in the row
    uint32_t my_vu = 0, run_reg = 0;
    uint64_t my_dst = 0;
    if (l < nd) {
      const TableDesc d = desc[tb + l];
      my_vu = d.val_u;
      if (WRITE) {
        my_dst = reinterpret_cast<uint64_t>(d.val + d.off[i]);
        run_reg = ccnt[c * ndmax + l];
      }
    }
    // two 64-entry groups per step: both groups' loads are in flight before
    // either group's ballots
    for (uint64_t t = t0; t < t1; t += 128) {
      const bool iv0 = t + l < t1, iv1 = t + 64 + l < t1;
      const uint32_t wv0 = iv0 ? gnbr[o0 + t + l] : 0;
      const uint32_t wv1 = iv1 ? gnbr[o0 + t + 64 + l] : 0;
      for (uint32_t j = 0; j < nd; ++j) {
        const uint32_t vu = readlane_u32(my_vu, j);
        const uint64_t *bm = bitmap + (uint64_t)vu * W;
        const uint64_t word0 = iv0 ? bm[wv0 >> 6] : 0, word1 = iv1 ? bm[wv1 >> 6] : 0;
        const bool hit0 = (word0 >> (wv0 & 63)) & 1ull, hit1 = (word1 >> (wv1 & 63)) & 1ull;
        const uint64_t m0 = __ballot(hit0), m1 = __ballot(hit1);
        const uint32_t run = readlane_u32(run_reg, j);
        if (WRITE && (hit0 || hit1)) {
          uint32_t *dst = reinterpret_cast<uint32_t *>(readlane_u64(my_dst, j));
          const uint32_t *rk = rank + (uint64_t)vu * (W + 1);
          if (hit0) dst[run + rank_below(m0)] = rk[wv0 >> 6] + (uint32_t)__popcll(word0 & ((1ull << (wv0 & 63)) - 1));
          if (hit1)
            dst[run + (uint32_t)__popcll(m0) + rank_below(m1)] =
                rk[wv1 >> 6] + (uint32_t)__popcll(word1 & ((1ull << (wv1 & 63)) - 1));
        }
        if (l == j) run_reg = run + (uint32_t)__popcll(m0) + (uint32_t)__popcll(m1);
      }
    }
    if (!WRITE && l < nd) ccnt[c * ndmax + l] = run_reg;
  }
}

// One thread per (heavy row, table slot): chunk counts -> chunk offsets inside
// the row, and the row's total as its count.
__global__ void k_theavy_bases(const TableJob job, const TableDesc *__restrict__ desc,
                               const uint64_t *__restrict__ heavy_list, const uint32_t *__restrict__ heavy_n,
                               const uint64_t *__restrict__ cpre, uint32_t *__restrict__ ccnt, uint32_t ndmax) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (uint64_t)(*heavy_n) * ndmax) return;
  const uint32_t h = (uint32_t)(x / ndmax), j = (uint32_t)(x % ndmax);
  const uint64_t e = heavy_list[h];
  const uint32_t ju = (uint32_t)(e >> 32), i = (uint32_t)e;
  if (j >= job.nd[ju]) return;
  uint32_t run = 0;
  for (uint64_t c = cpre[h]; c < cpre[h + 1]; ++c) {
    const uint32_t n = ccnt[c * ndmax + j];
    ccnt[c * ndmax + j] = run;
    run += n;
  }
  desc[job.tab0[ju] + j].off[i] = run;
}

void run_tables(gem_ctx *c) { run_tables(c, c->q); }

void run_tables(gem_ctx *c, QueryDev &Qd) {
  if (!Qd.filtered) throw Error{GEM_E_STATE, "filter has not run"};
  const gemh::QueryInfo &Q = Qd.Q;
  const GraphDev &G = graph(c);
  hipStream_t s = c->stream;
  GEM_HIP_CHECK(hipEventRecord(c->ev[2], s));
  Qd.tab.clear();
  Qd.tab.resize(2 * Q.ne);
  TableJob job{};
  job.W = Qd.W;
  std::vector<uint32_t> desc_tab;  // TableDesc index -> directed edge
  uint32_t waves = 0;
  for (uint32_t u = 0; u < Q.nv; ++u) {
    if (Q.deg[u] == 1) continue;  // leaves key no table (unifiedTrie.cu:1045,1095)
    const uint32_t ju = job.nu++;
    job.wave_start[ju] = waves;
    job.ncan[ju] = Qd.ncan[u];
    job.can_off[ju] = Qd.can_off[u];
    job.tab0[ju] = (uint32_t)desc_tab.size();
    for (auto [un, e] : Q.adj[u]) {
      const uint32_t d = 2 * e + (Q.edges[e].first == u ? 0 : 1);
      TableDev &T = Qd.tab[d];
      T.present = true;
      T.key_u = u;
      T.val_u = un;
      T.off.alloc(((uint64_t)Qd.ncan[u] + 1) * 8, s);
      desc_tab.push_back(d);
      job.nd[ju]++;
    }
    waves += (Qd.ncan[u] + 63) / 64;
  }
  job.wave_start[job.nu] = waves;
  std::vector<TableDesc> desc(desc_tab.size());
  auto fill_desc = [&]() {
    for (size_t k = 0; k < desc_tab.size(); ++k) {
      TableDev &T = Qd.tab[desc_tab[k]];
      desc[k] = TableDesc{T.val_u, T.off.as<uint64_t>(), T.val.as<uint32_t>()};
    }
  };
  fill_desc();
  DevBuf ddesc;
  ddesc.alloc(std::max<size_t>(desc.size(), 1) * sizeof(TableDesc), s);
  GEM_HIP_CHECK(hipMemcpyAsync(ddesc.p, desc.data(), desc.size() * sizeof(TableDesc), hipMemcpyHostToDevice, s));
  const uint32_t blocks = (waves + TAB_WAVES - 1) / TAB_WAVES;
  DevBuf adj_len, heavy_list, heavy_n;
  adj_len.alloc(GEM_MAX_QV * 8, s);
  GEM_HIP_CHECK(hipMemsetAsync(adj_len.p, 0, GEM_MAX_QV * 8, s));
  uint64_t rows_total = 0;
  for (uint32_t ju = 0; ju < job.nu; ++ju) rows_total += job.ncan[ju];
  heavy_list.alloc(std::max<uint64_t>(rows_total, 1) * 8, s);
  heavy_n.alloc(4, s);
  GEM_HIP_CHECK(hipMemsetAsync(heavy_n.p, 0, 4, s));
  uint32_t ndmax = 1;
  for (uint32_t ju = 0; ju < job.nu; ++ju) ndmax = std::max(ndmax, job.nd[ju]);
  DevBuf cpre, ccnt;
  cpre.alloc((std::max<uint64_t>(rows_total, 1) + 2) * 8, s);
  uint64_t *d_chunks = cpre.as<uint64_t>() + std::max<uint64_t>(rows_total, 1) + 1;
  uint32_t nh = 0;
  uint64_t nchunks = 0;
  auto launch_light = [&](bool write) {
    if (!blocks) return;
    if (write)
      k_tables<true><<<blocks, TAB_WAVES * 64, 0, s>>>(
          G.off.as<uint64_t>(), G.nbr_dev(), Qd.can.as<uint32_t>(), Qd.bitmap.as<uint64_t>(),
          Qd.rank.as<uint32_t>(), job, ddesc.as<TableDesc>(), adj_len.as<unsigned long long>(),
          heavy_list.as<uint64_t>(), heavy_n.as<uint32_t>());
    else
      k_tables<false><<<blocks, TAB_WAVES * 64, 0, s>>>(
          G.off.as<uint64_t>(), G.nbr_dev(), Qd.can.as<uint32_t>(), Qd.bitmap.as<uint64_t>(),
          Qd.rank.as<uint32_t>(), job, ddesc.as<TableDesc>(), adj_len.as<unsigned long long>(),
          heavy_list.as<uint64_t>(), heavy_n.as<uint32_t>());
    GEM_LAUNCH_CHECK(write ? "k_tables<write>" : "k_tables<count>");
  };
  auto launch_heavy = [&](bool write) {
    if (!nchunks) return;
    const uint32_t hb = (uint32_t)std::min<uint64_t>((nchunks + TAB_WAVES - 1) / TAB_WAVES, 2048);
    if (write)
      k_tables_heavy<true><<<hb, TAB_WAVES * 64, 0, s>>>(
          G.off.as<uint64_t>(), G.nbr_dev(), Qd.can.as<uint32_t>(), Qd.bitmap.as<uint64_t>(),
          Qd.rank.as<uint32_t>(), job, ddesc.as<TableDesc>(), adj_len.as<unsigned long long>(),
          heavy_list.as<uint64_t>(), heavy_n.as<uint32_t>(), cpre.as<uint64_t>(), ccnt.as<uint32_t>(), ndmax);
    else
      k_tables_heavy<false><<<hb, TAB_WAVES * 64, 0, s>>>(
          G.off.as<uint64_t>(), G.nbr_dev(), Qd.can.as<uint32_t>(), Qd.bitmap.as<uint64_t>(),
          Qd.rank.as<uint32_t>(), job, ddesc.as<TableDesc>(), adj_len.as<unsigned long long>(),
          heavy_list.as<uint64_t>(), heavy_n.as<uint32_t>(), cpre.as<uint64_t>(), ccnt.as<uint32_t>(), ndmax);
    GEM_LAUNCH_CHECK(write ? "k_tables_heavy<write>" : "k_tables_heavy<count>");
  };
  {
    ProfLaunch pl(c, K_TABLES_COUNT);
    launch_light(false);
    // rows longer than HEAVY_ROW: chunk prefix, then (sizes known) the chunked count
    k_theavy_prefix<<<1, 1024, 0, s>>>(G.off.as<uint64_t>(), Qd.can.as<uint32_t>(), job, heavy_list.as<uint64_t>(),
                                       heavy_n.as<uint32_t>(), cpre.as<uint64_t>(), d_chunks);
    GEM_LAUNCH_CHECK("k_theavy_prefix");
  }
  {
    uint64_t *hh = pinned<uint64_t>(c, 2);
    hh[0] = 0;
    GEM_HIP_CHECK(hipMemcpyAsync(hh, heavy_n.p, 4, hipMemcpyDeviceToHost, s));
    GEM_HIP_CHECK(hipMemcpyAsync(hh + 1, d_chunks, 8, hipMemcpyDeviceToHost, s));
    GEM_HIP_CHECK(hipStreamSynchronize(s));
    nh = (uint32_t)hh[0];
    nchunks = hh[1];
  }
  if (nchunks) {
    ccnt.alloc(nchunks * ndmax * 4, s);
    ProfLaunch pl(c, K_TABLES_COUNT);
    launch_heavy(false);
    const uint64_t nthreads = (uint64_t)nh * ndmax;
    k_theavy_bases<<<(uint32_t)((nthreads + 255) / 256), 256, 0, s>>>(
        job, ddesc.as<TableDesc>(), heavy_list.as<uint64_t>(), heavy_n.as<uint32_t>(), cpre.as<uint64_t>(),
        ccnt.as<uint32_t>(), ndmax);
    GEM_LAUNCH_CHECK("k_theavy_bases");
  }
  // row counts -> offsets for every table in one launch (the larger ones by
  // the multi-block scan), then the table sizes and adjacency lengths in one
  // copy back
  const size_t nt = desc_tab.size();
  std::vector<ScanSeg> segs(nt);
  for (size_t k = 0; k < nt; ++k) {
    TableDev &T = Qd.tab[desc_tab[k]];
    segs[k] = ScanSeg{T.off.as<uint64_t>(), (uint64_t)Qd.ncan[T.key_u]};
  }
  DevBuf dsegs, dtot;
  dsegs.alloc(std::max<size_t>(nt, 1) * sizeof(ScanSeg), s);
  dtot.alloc((nt + GEM_MAX_QV) * 8, s);
  if (nt) GEM_HIP_CHECK(hipMemcpyAsync(dsegs.p, segs.data(), nt * sizeof(ScanSeg), hipMemcpyHostToDevice, s));
  scan_many_u64(c, dsegs.as<ScanSeg>(), segs.data(), (uint32_t)nt, dtot.as<uint64_t>());
  GEM_HIP_CHECK(hipMemcpyAsync(dtot.as<uint64_t>() + nt, adj_len.p, GEM_MAX_QV * 8, hipMemcpyDeviceToDevice, s));
  uint64_t *h = pinned<uint64_t>(c, nt + GEM_MAX_QV);
  GEM_HIP_CHECK(hipMemcpyAsync(h, dtot.p, (nt + GEM_MAX_QV) * 8, hipMemcpyDeviceToHost, s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  // algorithmic bytes: per keyed u, its candidates (4 B), offset pairs (8 B)
  // and adjacency (4 B per slot) once for all its tables; per table the value
  // vertex's bitmap (|V|/8) and the row counts / offsets (8 B per row)
  double shared_bytes = 0, count_bytes = 0, write_bytes = 0;
  for (uint32_t ju = 0; ju < job.nu; ++ju) shared_bytes += 12.0 * job.ncan[ju] + 4.0 * (double)h[nt + ju];
  for (size_t k = 0; k < nt; ++k) {
    TableDev &T = Qd.tab[desc_tab[k]];
    T.size = h[k];
    count_bytes += 8.0 * Qd.W + 8.0 * Qd.ncan[T.key_u];
    write_bytes += 12.0 * Qd.W + 8.0 * Qd.ncan[T.key_u] + 4.0 * (double)T.size;
    T.val.alloc(std::max<uint64_t>(T.size, 1) * 4, s);
  }
  prof_bytes(c, K_TABLES_COUNT, shared_bytes + count_bytes);
  prof_bytes(c, K_TABLES_WRITE, shared_bytes + write_bytes);
  fill_desc();
  GEM_HIP_CHECK(hipMemcpyAsync(ddesc.p, desc.data(), desc.size() * sizeof(TableDesc), hipMemcpyHostToDevice, s));
  {
    ProfLaunch pl(c, K_TABLES_WRITE);
    launch_light(true);
    launch_heavy(true);
  }
  GEM_HIP_CHECK(hipEventRecord(c->ev[3], s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  prof_flush(c);
  float ms = 0;
  GEM_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  c->times.tables_ms = ms;
  Qd.tabled = true;
}

}  // namespace gemd
