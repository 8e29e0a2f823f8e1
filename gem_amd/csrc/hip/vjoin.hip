// Vertex-at-a-time join over many queries at once (the batched path).
//
// Same counts as the edge-at-a-time join (the reference's join(),
// in-core/lib/join/src/join_trie.cu:1087-1277), which is order-invariant, but
// each level adds one query vertex and applies ALL its edges to earlier
// vertices at once: the candidates of the new vertex are the intersection of
// the backward tables' rows (sorted position lists in its candidate set), so a
// closing edge prunes rows before they are ever written instead of after a
// whole level has been materialised.
//
//   k_vlen    per 64-row unit: sum over rows of the shortest list length
//             (an upper bound on the level's output, used to size and split)
//   k_vext    waves walk the shortest ("driver") list of each of their 64
//             rows, 128 items per step (two per lane); each item is
//             binary-searched in the row's other lists and compared with the
//             equal-label columns; survivors are counted (last level) or
//             appended to the output (one atomic reservation per wave step;
//             row order is free); rows with long driving lists are left to
//             k_vext_heavy
//   k_vext_heavy  walks those rows in fixed-size pieces spread over waves
//   k_vcount1 last level with a single backward list: |list| minus the
//             equal-label columns found in it (no walk)
//
// Work distribution: each launch covers many (query, level, row range) items;
// waves pull 64-row tasks from a device counter (persistent grid), so queries
// of very different sizes share one launch.  The host scheduler
// (host/vsched.h) keeps, per query, a depth-first stack of pending levels and
// cuts them into 64-row-unit ranges whose output bounds fit a memory budget
// shared by all queries; a query stops as soon as its count exceeds the cap.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "ctx.h"
#include "tcache.h"
#include "vjoin.h"
#include "wave.h"

namespace gemd {

constexpr uint32_t VB = 256;  // rows per task block
constexpr uint32_t VW = VB / 64;
// survivors a wave stages in LDS before it reserves output rows for them (a
// step adds at most 128)
constexpr uint32_t VSTAGE = 512;

__device__ __forceinline__ uint32_t vposition_in(const uint64_t *bm, const uint32_t *rank, uint32_t v) {
  const uint64_t word = ldg(bm + (v >> 6));
  const uint32_t b = v & 63;
  if (!((word >> b) & 1ull)) return NOT_FOUND;
  return ldg(rank + (v >> 6)) + (uint32_t)__popcll(word & ((1ull << b) - 1));
}

// A level descriptor read through the global address space (see ldg).
using GLevel = const __attribute__((address_space(1))) VLevel;

// task g (global block index of the launch) -> item t, row block within item
__device__ __forceinline__ uint32_t find_item(const uint64_t *blk_prefix, uint32_t nitems, uint64_t g) {
  uint32_t lo = 0, hi = nitems;  // largest t with blk_prefix[t] <= g
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (blk_prefix[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------ length pass
// Per 64-row unit (wave w of 256-row block g -> unit_u[4g + w]): the sum over
// its rows of the shortest backward list, an upper bound on its output rows.
__global__ void __launch_bounds__(VB) k_vlen(const VItem *__restrict__ items, uint32_t nitems,
                                            const uint64_t *__restrict__ blk_prefix,
                                            unsigned int *__restrict__ queue, uint64_t *__restrict__ unit_u) {
  __shared__ uint32_t task;
  const uint64_t total = blk_prefix[nitems];
  for (;;) {
    if (threadIdx.x == 0) task = atomicAdd(queue, 1u);
    __syncthreads();
    const uint64_t g = task;
    __syncthreads();
    if (g >= total) break;
    const uint32_t t = find_item(blk_prefix, nitems, g);
    const VItem it = items[t];
    GLevel &L = *(GLevel *)it.lv;
    const uint64_t r = it.row_begin + (g - blk_prefix[t]) * VB + threadIdx.x;
    uint64_t u = 0;
    if (r < it.row_end) {
      const uint32_t *row = it.rows + r * it.k;
      u = ~0ull;
      for (uint32_t j = 0; j < L.nb; ++j) {
        const uint32_t p = ldg(row + L.key_col[j]);
        const uint64_t *off = L.toff[j];
        const uint64_t len = ldg(off + p + 1) - ldg(off + p);
        u = len < u ? len : u;
      }
    }
    u = wave_sum_u64(u);
    if (threadIdx.x % 64 == 0) unit_u[g * VW + threadIdx.x / 64] = u;
  }
}

// Device address of the list of key p in backward table j (b = its first
// offset): in HBM, or through the streamed tables' block map.
template <bool OOC>
__device__ __forceinline__ uint64_t list_addr(GLevel &L, uint32_t j, uint32_t p, uint64_t b) {
  if (OOC) return ldg(L.vmap[j] + (p >> L.kshift[j])) + 4 * b;
  return (uint64_t)(uintptr_t)(L.tval[j] + b);
}

// ------------------------------------------------------ extend (walk) pass
// Each wave pulls 64-row tasks from the launch's queue on its own (no block
// barrier).  For each of its rows the shortest backward list drives; a row
// whose driving list is longer than heavy_t is not walked here but appended to
// the heavy-row list (when it has room), which k_vext_heavy then walks in
// fixed-size chunks spread over all waves: a hub row no longer holds one wave
// for the whole launch.
// Dynamic LDS per wave: other lists (nb-1) x 64 x (u64 address + u32 len),
// dup positions ndup x 64 u32 (blocks of one wave when four waves' share
// would not fit a block's LDS).
// Survivors are written as whole rows: their (parent row, new value) pairs are
// staged in LDS, then the wave writes the step's new rows as one run of
// consecutive words.
// The item count is read from device memory so that one captured graph can
// replay the launch for any number of items.  Tasks [g_begin, g_end) of the
// launch (256-row units) are processed (streamed tables run a launch in task
// groups).
template <bool FINAL, bool OOC>
__global__ void __launch_bounds__(VB) __attribute__((amdgpu_waves_per_eu(8)))
k_vext(const VItem *__restrict__ items, const uint32_t *__restrict__ nitems_p, const uint64_t *__restrict__ blk_prefix,
       unsigned int *__restrict__ queue, uint32_t max_nb, uint32_t max_ndup,
       unsigned long long *__restrict__ prof_items, unsigned long long *__restrict__ prof_out, uint64_t g_begin,
       uint64_t g_end, HeavyRow *__restrict__ heavy, unsigned int *__restrict__ heavy_n, uint32_t heavy_cap,
       uint32_t heavy_t) {
  const uint32_t nitems = *nitems_p;
  __shared__ uint32_t st[VW][65];
  __shared__ uint64_t drv[VW][64];  // address of each row's driving list
  __shared__ uint8_t sv_row[VW][VSTAGE];    // staged survivors of the task: row in the task
  __shared__ uint32_t sv_val[VW][VSTAGE];  // and new value
  extern __shared__ unsigned char vlds[];
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t per_wave = 64 * ((max_nb > 0 ? max_nb - 1 : 0) * 12 + max_ndup * 4);
  unsigned char *mine = vlds + (size_t)w * per_wave;
  uint64_t *obase = reinterpret_cast<uint64_t *>(mine);                                  // [nb-1][64]
  uint32_t *olen = reinterpret_cast<uint32_t *>(mine + (size_t)64 * 8 * (max_nb - 1));   // [nb-1][64]
  uint32_t *dpos = reinterpret_cast<uint32_t *>(mine + (size_t)64 * 12 * (max_nb - 1));  // [ndup][64]
  const uint64_t total = blk_prefix[nitems] < g_end ? blk_prefix[nitems] : g_end;
  uint64_t walked = 0, written = 0;
  for (;;) {
    uint32_t wt = 0;
    if (l == 0) wt = atomicAdd(queue, 1u);
    wt = readlane_u32(wt, 0);
    const uint64_t q = g_begin * VW + wt;  // 64-row task
    if (q >= total * VW) break;
    const uint64_t g = q / VW;
    const uint32_t t = find_item(blk_prefix, nitems, g);
    const VItem it = items[t];
    GLevel &L = *(GLevel *)it.lv;
    const uint32_t nb = L.nb, ndup = L.ndup, k = it.k;
    const uint64_t rbase = it.row_begin + (g - blk_prefix[t]) * VB + (q % VW) * 64;
    const uint64_t r = rbase + l;
    const bool valid = r < it.row_end;
    const uint32_t *row = it.rows + r * k;
    // per-row lists: the shortest one drives, the others are probed
    uint32_t dlen = 0, dj = 0, dp = 0;
    uint64_t dbase = 0, daddr = 0;
    if (valid) {
      dlen = 0xFFFFFFFFu;
      for (uint32_t j = 0; j < nb; ++j) {
        const uint32_t p = ldg(row + L.key_col[j]);
        const uint64_t *off = L.toff[j];
        const uint64_t b = ldg(off + p);
        const uint32_t len = (uint32_t)(ldg(off + p + 1) - b);
        if (len < dlen) {
          dlen = len;
          dj = j;
          dbase = b;
          dp = p;
        }
      }
      if (heavy && dlen > heavy_t) {
        const uint32_t h = atomicAdd(heavy_n, 1u);
        if (h < heavy_cap) {
          heavy[h] = HeavyRow{t, dlen, r};
          dlen = 0;  // walked by k_vext_heavy
        }
      }
    }
    // a row with nothing to walk here (empty driving list, or set aside as a
    // heavy row) owns no item of the steps below, so its probe lists and
    // equal-label positions are never read: skip their loads
    if (valid && dlen) {
      daddr = list_addr<OOC>(L, dj, dp, dbase);
      uint32_t o = 0;
      for (uint32_t j = 0; j < nb; ++j) {
        if (j == dj) continue;
        const uint32_t p = ldg(row + L.key_col[j]);
        const uint64_t *off = L.toff[j];
        const uint64_t b = ldg(off + p);
        obase[o * 64 + l] = list_addr<OOC>(L, j, p, b);
        olen[o * 64 + l] = (uint32_t)(ldg(off + p + 1) - b);
        ++o;
      }
      for (uint32_t c = 0; c < ndup; ++c) {
        const uint32_t vid = ldg(L.dup_can[c] + ldg(row + L.dup_col[c]));
        dpos[c * 64 + l] = vposition_in(L.bm_new, L.rank_new, vid);
      }
    }
    const uint32_t incl = wave_incl_scan_u32(dlen);
    const uint32_t wtotal = readlane_u32(incl, 63);
    st[w][l] = incl - dlen;
    if (l == 63) st[w][64] = wtotal;
    drv[w][l] = daddr;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    walked += wtotal;
    uint64_t survivors = 0;
    uint32_t pend = 0;  // survivors staged in LDS, not yet written
    // write the staged survivors: one reservation for all of them, then
    // their rows as one run of consecutive words (word q is column q % (k+1)
    // of survivor q / (k+1): k parent columns, then the value)
    auto flush = [&]() {
      uint64_t base = 0;
      if (l == 0) base = atomicAdd(it.counter, (unsigned long long)pend);
      base = readlane_u64(base, 0);
      written += (uint64_t)pend * 4 * (k + 1);
      const uint32_t rl = k + 1;
      for (uint32_t q = l; q < pend * rl; q += 64) {
        const uint32_t i = q / rl, c = q - i * rl;
        if (base + i < it.out_cap)
          stg(it.out + (base + i) * rl + c, c < k ? ldg(it.rows + (rbase + sv_row[w][i]) * k + c) : sv_val[w][i]);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      pend = 0;
    };
    // 256 items per step, four per lane (items s + 64 g + l), so that each
    // lane has four loads in flight at every stage of the probes
    for (uint32_t s = 0; s < wtotal; s += 256) {
      uint32_t val[4], rw[4];
      bool keep[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const uint32_t item = s + 64u * g + l;
        keep[g] = item < wtotal;
        rw[g] = keep[g] ? seg_of(st[w], item) : 0;
        val[g] = keep[g] ? ldg(reinterpret_cast<const uint32_t *>(drv[w][rw[g]]) + (item - st[w][rw[g]])) : 0;
      }
      for (uint32_t c = 0; c < ndup; ++c) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
          if (dpos[c * 64 + rw[g]] == val[g]) keep[g] = false;
      }
      for (uint32_t o = 0; o + 1 < nb; ++o) {
        const uint32_t *lists[4];
        uint32_t lens[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          lists[g] = reinterpret_cast<const uint32_t *>(obase[o * 64 + rw[g]]);
          lens[g] = olen[o * 64 + rw[g]];
        }
        contains_quad(lists, lens, val, keep);
      }
      uint64_t m[4];
      uint32_t nk = 0, before[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        m[g] = __ballot(keep[g]);
        before[g] = nk;
        nk += (uint32_t)__popcll(m[g]);
      }
      if (FINAL) {
        survivors += nk;
      } else if (nk) {
        // stage this step's survivors after the pending ones; the staging
        // area is written out when the next step might not fit
#pragma unroll
        for (int g = 0; g < 4; ++g)
          if (keep[g]) {
            const uint32_t i = pend + before[g] + rank_below(m[g]);
            sv_row[w][i] = rw[g];
            sv_val[w][i] = val[g];
          }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        pend += nk;
        if (pend > VSTAGE - 256) flush();
      }
    }
    if (!FINAL && pend) flush();
    if (FINAL && l == 0 && survivors) atomicAdd(it.counter, (unsigned long long)survivors);
    // st / drv / LDS lists of this wave are rewritten by its next task
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (prof_items && l == 0 && walked) atomicAdd(prof_items, (unsigned long long)walked);
  if (prof_out && l == 0 && written) atomicAdd(prof_out, (unsigned long long)written);
}

// One block: chunk prefix of the heavy rows, pre[i] = sum over rows before i
// of ceil(len / chunk), pre[n] = total, n = min(heavy_n, heavy_cap).
__global__ void __launch_bounds__(1024) k_heavy_scan(const HeavyRow *__restrict__ heavy,
                                                    const unsigned int *__restrict__ heavy_n, uint32_t heavy_cap,
                                                    uint32_t chunk, uint64_t *__restrict__ pre) {
  __shared__ uint64_t sm[16];
  const uint32_t n = *heavy_n < heavy_cap ? *heavy_n : heavy_cap;
  uint64_t carry = 0;
  for (uint32_t b = 0; b < n; b += blockDim.x) {
    const uint32_t i = b + threadIdx.x;
    const uint64_t x = i < n ? (heavy[i].len + chunk - 1) / chunk : 0;
    uint64_t ex;
    const uint64_t tot = block_excl_scan_u64(x, &ex, sm);
    if (i < n) pre[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) pre[n] = carry;
}

// Entries of a staged window of the first other list (k_vext_heavy).
constexpr uint32_t HWIN = 512;

// Heavy rows: each wave task walks `chunk` entries of one row's driving list
// (the same list the main pass picked: shortest, first on ties), probing the
// row's other lists and equal-label columns; survivors are counted or appended
// to the row's item output like the main pass's.
template <bool FINAL, bool OOC>
__global__ void __launch_bounds__(VB) __attribute__((amdgpu_waves_per_eu(8)))
k_vext_heavy(const VItem *__restrict__ items, const HeavyRow *__restrict__ heavy,
             const unsigned int *__restrict__ heavy_n, uint32_t heavy_cap, const uint64_t *__restrict__ pre,
             unsigned int *__restrict__ queue, uint32_t chunk, unsigned long long *__restrict__ prof_items,
             unsigned long long *__restrict__ prof_out) {
  __shared__ uint64_t oaddr[VW][GEM_MAX_QV];
  __shared__ uint32_t olen[VW][GEM_MAX_QV], dpos[VW][GEM_MAX_QV];
  __shared__ uint32_t wend[VW][64];   // window ends in the first other list
  __shared__ uint32_t win[VW][HWIN];  // staged window
  __shared__ uint32_t prow[VW][GEM_MAX_QV], hsv[VW][128];  // parent row, this step's surviving values
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t n = *heavy_n < heavy_cap ? *heavy_n : heavy_cap;
  const uint64_t total = pre[n];
  uint64_t walked = 0, written = 0;
  for (;;) {
    uint32_t wt = 0;
    if (l == 0) wt = atomicAdd(queue, 1u);
    wt = readlane_u32(wt, 0);
    if (wt >= total) break;
    const uint32_t h = find_item(pre, n, wt);
    const HeavyRow e = heavy[h];
    const VItem it = items[e.t];
    GLevel &L = *(GLevel *)it.lv;
    const uint32_t nb = L.nb, ndup = L.ndup, k = it.k;
    const uint32_t *row = it.rows + e.r * k;
    // the row's lists (every lane computes the same; lane j keeps list j)
    uint32_t dlen = 0xFFFFFFFFu, dj = 0, dp = 0;
    uint64_t dbase = 0;
    for (uint32_t j = 0; j < nb; ++j) {
      const uint32_t p = ldg(row + L.key_col[j]);
      const uint64_t *off = L.toff[j];
      const uint64_t b = ldg(off + p);
      const uint32_t len = (uint32_t)(ldg(off + p + 1) - b);
      if (len < dlen) {
        dlen = len;
        dj = j;
        dbase = b;
        dp = p;
      }
    }
    const uint32_t *dlist = reinterpret_cast<const uint32_t *>(list_addr<OOC>(L, dj, dp, dbase));
    if (l < nb) {
      const uint32_t j = l;
      const uint32_t p = ldg(row + L.key_col[j]);
      const uint64_t b = ldg(L.toff[j] + p);
      const uint32_t o = j < dj ? j : j - 1;  // other lists, in order, without the driver
      if (j != dj) {
        oaddr[w][o] = list_addr<OOC>(L, j, p, b);
        olen[w][o] = (uint32_t)(ldg(L.toff[j] + p + 1) - b);
      }
    }
    if (l < ndup) dpos[w][l] = vposition_in(L.bm_new, L.rank_new, ldg(L.dup_can[l] + ldg(row + L.dup_col[l])));
    if (!FINAL && l < k) prow[w][l] = ldg(row + l);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint64_t s0 = (uint64_t)(wt - pre[h]) * chunk;
    const uint64_t s1 = s0 + chunk < dlen ? s0 + chunk : dlen;
    walked += s1 - s0;
    uint64_t survivors = 0;
    // Two items per lane per step, as in k_vext.  The first other list is
    // probed through windows: for each 128-entry step of the chunk, the range
    // of that list that can hold the step's values (both lists are sorted and
    // duplicate-free) is found for up to 63 steps at once (lane j: the end of
    // step j, lane 63: the start of the first step), and a window of at most
    // HWIN entries is staged in LDS with coalesced loads and searched there.
    const uint32_t *l0p = nb > 1 ? reinterpret_cast<const uint32_t *>(oaddr[w][0]) : nullptr;
    const uint32_t l0n = nb > 1 ? olen[w][0] : 0;
    for (uint64_t S = s0; S < s1; S += 63 * 128) {
      const uint64_t S1 = S + 63 * 128 < s1 ? S + 63 * 128 : s1;
      const uint32_t nsub = (uint32_t)((S1 - S + 127) / 128);
      if (nb > 1) {
        uint32_t pos = 0;
        if (l < nsub) {
          const uint64_t last = (S + 128ull * (l + 1) < S1 ? S + 128ull * (l + 1) : S1) - 1;
          pos = upper_bound_u32(l0p, l0n, ldg(dlist + last));
        } else if (l == 63) {
          pos = lower_bound_u32(l0p, l0n, ldg(dlist + S));
        }
        wend[w][l] = pos;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      for (uint32_t j = 0; j < nsub; ++j) {
        const uint64_t s = S + 128ull * j;
        bool keep0 = s + l < S1, keep1 = s + 64 + l < S1;
        uint32_t val0 = 0, val1 = 0;
        if (keep0) val0 = ldg(dlist + s + l);
        if (keep1) val1 = ldg(dlist + s + 64 + l);
        for (uint32_t c = 0; c < ndup; ++c) {
          if (dpos[w][c] == val0) keep0 = false;
          if (dpos[w][c] == val1) keep1 = false;
        }
        if (nb > 1) {
          const uint32_t a = j ? wend[w][j - 1] : wend[w][63], wn = wend[w][j] - a;
          if (wn <= HWIN) {
            for (uint32_t i = l; i < wn; i += 64) win[w][i] = ldg(l0p + a + i);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            contains_pair<false>(win[w], wn, val0, keep0, win[w], wn, val1, keep1);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          } else {
            contains_pair(l0p + a, wn, val0, keep0, l0p + a, wn, val1, keep1);
          }
        }
        for (uint32_t o = 1; o + 1 < nb; ++o) {
          const uint32_t *a = reinterpret_cast<const uint32_t *>(oaddr[w][o]);
          contains_pair(a, olen[w][o], val0, keep0, a, olen[w][o], val1, keep1);
        }
        const uint64_t m0 = __ballot(keep0), m1 = __ballot(keep1);
        const uint32_t nk0 = (uint32_t)__popcll(m0), nk = nk0 + (uint32_t)__popcll(m1);
        if (FINAL) {
          survivors += nk;
        } else if (nk) {
          uint64_t base = 0;
          if (l == 0) base = atomicAdd(it.counter, (unsigned long long)nk);
          base = readlane_u64(base, 0);
          written += (uint64_t)nk * 4 * (k + 1);
          // the nk new rows share the parent row: they are nk * (k + 1)
          // consecutive words of the output, written by consecutive lanes
          if (keep0) hsv[w][rank_below(m0)] = val0;
          if (keep1) hsv[w][nk0 + rank_below(m1)] = val1;
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          const uint32_t rl = k + 1;
          for (uint32_t q = l; q < nk * rl; q += 64) {
            const uint32_t r = q / rl, c = q - r * rl;
            if (base + r < it.out_cap) stg(it.out + (base + r) * rl + c, c < k ? prow[w][c] : hsv[w][r]);
          }
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      // wend of this wave is rewritten by the next group of steps
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (FINAL && l == 0 && survivors) atomicAdd(it.counter, (unsigned long long)survivors);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (prof_items && l == 0 && walked) atomicAdd(prof_items, (unsigned long long)walked);
  if (prof_out && l == 0 && written) atomicAdd(prof_out, (unsigned long long)written);
}

// Last level, one backward list: |list| - equal-label columns present.
template <bool OOC>
__global__ void __launch_bounds__(VB) k_vcount1(const VItem *__restrict__ items, const uint32_t *__restrict__ nitems_p,
                                               const uint64_t *__restrict__ blk_prefix,
                                               unsigned int *__restrict__ queue, uint64_t g_begin, uint64_t g_end) {
  const uint32_t nitems = *nitems_p;
  __shared__ uint32_t task;
  __shared__ uint64_t sm[VW];
  const uint64_t total = blk_prefix[nitems] < g_end ? blk_prefix[nitems] : g_end;
  for (;;) {
    if (threadIdx.x == 0) task = atomicAdd(queue, 1u);
    __syncthreads();
    const uint64_t g = g_begin + task;
    __syncthreads();
    if (g >= total) break;
    const uint32_t t = find_item(blk_prefix, nitems, g);
    const VItem it = items[t];
    GLevel &L = *(GLevel *)it.lv;
    const uint64_t r = it.row_begin + (g - blk_prefix[t]) * VB + threadIdx.x;
    uint64_t cnt = 0;
    if (r < it.row_end) {
      const uint32_t *row = it.rows + r * it.k;
      const uint32_t p = ldg(row + L.key_col[0]);
      const uint64_t b = ldg(L.toff[0] + p), e = ldg(L.toff[0] + p + 1);
      cnt = e - b;
      const uint32_t *lst = nullptr;
      if (L.ndup) lst = reinterpret_cast<const uint32_t *>(list_addr<OOC>(L, 0, p, b));
      for (uint32_t c = 0; c < L.ndup; ++c) {
        const uint32_t pos = vposition_in(L.bm_new, L.rank_new, ldg(L.dup_can[c] + ldg(row + L.dup_col[c])));
        if (pos != NOT_FOUND && sorted_contains(lst, e - b, pos)) --cnt;
      }
    }
    cnt = wave_sum_u64(cnt);
    if (threadIdx.x % 64 == 0) sm[threadIdx.x / 64] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t s = sm[0] + sm[1] + sm[2] + sm[3];
      if (s) atomicAdd(it.counter, (unsigned long long)s);
    }
    __syncthreads();
  }
}

// Streamed tables: for every table block a row of the launch will read, the
// first and last task (launch task + base) that reads it.  Blocks (256-row
// tasks) are strided over the grid.
// Only the launch's tasks g with base + g in [wa, wb) are marked, as
// base + g - wa (the iteration is planned one window of tasks at a time).
__global__ void __launch_bounds__(VB) k_vmark(const VItem *__restrict__ items, uint32_t nitems,
                                             const uint64_t *__restrict__ blk_prefix, uint64_t base, uint64_t wa,
                                             uint64_t wb, bool count_only, uint32_t *__restrict__ first,
                                             uint32_t *__restrict__ last) {
  const uint64_t total = blk_prefix[nitems];
  const uint64_t g0 = wa > base ? wa - base : 0;
  const uint64_t g1 = wb - base < total ? wb - base : total;
  for (uint64_t g = g0 + blockIdx.x; g < g1; g += gridDim.x) {
    const uint32_t t = find_item(blk_prefix, nitems, g);
    const VItem it = items[t];
    GLevel &L = *(GLevel *)it.lv;
    const uint64_t r = it.row_begin + (g - blk_prefix[t]) * VB + threadIdx.x;
    // k_vcount1 reads list values only to remove equal-label columns
    if (r >= it.row_end || (count_only && L.ndup == 0)) continue;
    const uint32_t *row = it.rows + r * it.k;
    const uint32_t gt = (uint32_t)(base + g - wa);
    for (uint32_t j = 0; j < L.nb; ++j) {
      const uint32_t b = L.gblk[j] + (ldg(row + L.key_col[j]) >> L.kshift[j]);
      if (first[b] > gt) atomicMin(&first[b], gt);
      if (last[b] < gt) atomicMax(&last[b], gt);
    }
  }
}

__global__ void k_viota(uint32_t *p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

// ================================================================== host

void build_vlevels(gem_ctx *c, VQueryState &S) {
  const QueryDev &Qd = *S.qd;
  const gemh::VertexPlan &P = S.plan;
  const uint32_t nv = Qd.Q.nv;
  std::vector<VLevel> lv(nv);
  for (uint32_t i = 1; i < nv; ++i) {
    const gemh::VLevelPlan &lp = P.levels[i];
    VLevel &L = lv[i];
    L.nb = lp.nb;
    L.ndup = lp.ndup;
    for (uint32_t j = 0; j < lp.nb; ++j) {
      const TableDev &T = Qd.tab[lp.directed[j]];
      if (!T.present) throw Error{GEM_E_STATE, "vertex plan uses a table that was not built"};
      L.key_col[j] = lp.key_col[j];
      L.toff[j] = T.off.as<uint64_t>();
      L.tval[j] = T.val.as<uint32_t>();
    }
    const uint32_t u = P.vorder[i];
    for (uint32_t d = 0; d < lp.ndup; ++d) {
      L.dup_col[d] = lp.dup_col[d];
      L.dup_can[d] = Qd.can.as<uint32_t>() + Qd.can_off[P.vorder[lp.dup_col[d]]];
    }
    L.bm_new = Qd.bitmap.as<uint64_t>() + (uint64_t)u * Qd.W;
    L.rank_new = Qd.rank.as<uint32_t>() + (uint64_t)u * (Qd.W + 1);
  }
  S.levels.alloc(sizeof(VLevel) * nv, c->stream);
  GEM_HIP_CHECK(hipMemcpyAsync(S.levels.p, lv.data(), sizeof(VLevel) * nv, hipMemcpyHostToDevice, c->stream));
  S.max_nb.assign(nv, 0);
  S.max_ndup.assign(nv, 0);
  for (uint32_t i = 1; i < nv; ++i) {
    S.max_nb[i] = lv[i].nb;
    S.max_ndup[i] = lv[i].ndup;
  }
  // the copy above reads lv asynchronously: keep it in the query state
  S.hlv = std::move(lv);
}

void replan(gem_ctx *c, VQueryState &S, const std::vector<uint32_t> &vorder) {
  if (gemh::make_vertex_plan(S.Qp ? *S.Qp : S.qd->Q, vorder, &S.plan))
    throw Error{GEM_E_INVALID, "invalid vertex order (leaf first, or a vertex without an earlier neighbour)"};
  if (S.edge_map)
    for (size_t l = 1; l < S.plan.levels.size(); ++l) {
      gemh::VLevelPlan &lp = S.plan.levels[l];
      for (uint32_t j = 0; j < lp.nb; ++j)
        lp.directed[j] = 2 * (*S.edge_map)[lp.directed[j] >> 1] + (lp.directed[j] & 1);
    }
  build_vlevels(c, S);
  S.mnb_all = std::max(S.mnb_all, *std::max_element(S.max_nb.begin(), S.max_nb.end()));
  S.mnd_all = std::max(S.mnd_all, *std::max_element(S.max_ndup.begin(), S.max_ndup.end()));
}

uint64_t attempt_limit(const VQueryState &S, uint32_t a) {
  if (!S.first_attempt || S.ranked.size() < 2) return 0;
  if (S.max_attempts && a + 1 >= S.max_attempts) return 0;
  const uint32_t rounds = a / std::max<uint32_t>(S.width, 1);
  const double lim = (double)S.first_attempt * std::pow((double)std::max<uint32_t>(S.growth, 2), (double)rounds);
  if (lim >= 9.0e18) return 0;  // beyond 2^63: no limit
  return (uint64_t)lim;
}

// (Re)start a query's attempt: the identity level of its first vertex's
// candidates, no count yet.
static void seed_query(gem_ctx *c, VQueryState &V) {
  hipStream_t s = c->stream;
  gemh::SQuery<DevBuf> &Q = V.s;
  Q.nb = V.max_nb;
  Q.count = 0;
  Q.fin_bound = Q.fin_got = 0;
  Q.stack.clear();
  Q.level_rows.assign(Q.nv, 0);
  const uint64_t n0 = V.qd->ncan[V.plan.vorder[0]];
  Q.level_rows[0] = n0;
  if (n0 == 0) {
    Q.done = true;
    return;
  }
  auto buf = std::make_shared<DevBuf>();
  buf->alloc(n0 * 4, s);
  k_viota<<<(uint32_t)((n0 + 255) / 256), 256, 0, s>>>(buf->as<uint32_t>(), n0);
  GEM_LAUNCH_CHECK("k_viota");
  gemh::SLevel<DevBuf> L;
  L.level = 1;
  L.rows = std::move(buf);
  L.nrows = n0;
  L.k = 1;
  Q.stack.push_back(std::move(L));
}

namespace {

struct Launch {
  std::vector<VItem> items;
  std::vector<uint64_t> blk;  // prefix, items.size() + 1
  uint32_t max_nb = 1, max_ndup = 0;
  void add(const VItem &it, uint32_t nb, uint32_t ndup) {
    if (blk.empty()) blk.push_back(0);
    items.push_back(it);
    blk.push_back(blk.back() + (it.row_end - it.row_begin + VB - 1) / VB);
    max_nb = std::max(max_nb, nb);
    max_ndup = std::max(max_ndup, ndup);
  }
  bool empty() const { return items.empty(); }
};

// Upload an item table + block prefix, return device pointers (kept alive by bufs).
struct DevLaunch {
  DevBuf items, blk, queue;
};

void upload(gem_ctx *c, const Launch &L, DevLaunch &D) {
  hipStream_t s = c->stream;
  D.items.alloc(sizeof(VItem) * L.items.size(), s);
  D.blk.alloc(8 * L.blk.size(), s);
  D.queue.alloc(4, s);
  GEM_HIP_CHECK(hipMemcpyAsync(D.items.p, L.items.data(), sizeof(VItem) * L.items.size(), hipMemcpyHostToDevice, s));
  GEM_HIP_CHECK(hipMemcpyAsync(D.blk.p, L.blk.data(), 8 * L.blk.size(), hipMemcpyHostToDevice, s));
  GEM_HIP_CHECK(hipMemsetAsync(D.queue.p, 0, 4, s));
}

// persistent grid: enough blocks to fill the chip (tunable cap), never more
// than the tasks
uint32_t g_persistent_blocks = 2048;
uint32_t grid_for(uint64_t blocks) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(blocks, g_persistent_blocks));
}

// Rows whose driving list has more entries than the context's heavy_t are
// walked by k_vext_heavy, heavy_chunk entries per wave task; at most HEAVY_CAP
// such rows per launch (a row that finds the list full is walked by its own
// wave).
constexpr uint32_t HEAVY_CAP = 1u << 20;

// Fixed buffers for the extend step of every iteration of one run: item
// tables and block prefixes of the three launches (write, final walk, final
// single-list count), the heavy-row lists of the two walking launches, and a
// header with the item counts, work queues, heavy-row counts, walked-entry
// counts and per-task result counters.  Host side: pinned mirrors.  Because
// every pointer, grid and LDS size is fixed for the run, the step
//   H2D header, items, prefixes -> 3 kernels -> 2 x (heavy scan, heavy walk)
//   -> D2H header
// is captured once as a hipGraph and replayed each iteration (no memset node:
// the zeroed counters travel in the header copy).
struct ExtSlot {
  // header: item counts [3], queues [3 main + 2 heavy], heavy-row counts [2],
  // profile counters [6] (u64: entries walked by the write / final / heavy
  // write / heavy final kernels, bytes written by the write / heavy write
  // kernels), result counters [cap] (u64)
  static constexpr size_t HDR_NITEMS = 0, HDR_QUEUE = 16, HDR_HEAVYN = 40, HDR_PITEMS = 48, HDR_CTR = 96;
  uint32_t cap = 0;  // items per launch = counters
  DevBuf items[3], blk[3], hdr;
  DevBuf heavy[2], hpre[2];  // heavy rows and their chunk prefix, per walking launch
  VItem *h_items[3] = {};
  uint64_t *h_blk[3] = {};
  unsigned char *h_hdr = nullptr;
  size_t hdr_bytes = 0;
  uint32_t max_nb = 1, max_ndup = 0;
  size_t lds = 0;
  uint32_t vthreads = VB;  // k_vext block size: 64 when four waves' lists do not fit one block's LDS
  uint32_t grid = 0;
  uint32_t heavy_t = 1024, heavy_chunk = 1024;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void init(gem_ctx *c, uint32_t capacity, uint32_t mnb, uint32_t mnd) {
    hipStream_t s = c->stream;
    cap = std::max<uint32_t>(capacity, 1);
    max_nb = std::max<uint32_t>(mnb, 1);
    max_ndup = mnd;
    // per wave: (nb - 1) list addresses and lengths and ndup equal-label
    // positions for each of its 64 rows (at most 31 KB with 32 query vertices)
    const size_t lds_wave = (size_t)64 * ((max_nb - 1) * 12 + max_ndup * 4);
    vthreads = lds_wave * VW <= 60 * 1024 ? VB : 64;
    lds = lds_wave * (vthreads / 64);
    if (lds > 60 * 1024) throw Error{GEM_E_LIMIT, "a level has too many backward edges for the vjoin LDS layout"};
    grid = g_persistent_blocks;
    heavy_t = c->heavy_t;
    heavy_chunk = c->heavy_chunk;
    hdr_bytes = HDR_CTR + 8ull * cap;
    hdr.alloc(hdr_bytes, s);
    for (int i = 0; i < 2; ++i) {
      heavy[i].alloc(sizeof(HeavyRow) * HEAVY_CAP, s);
      hpre[i].alloc(8ull * (HEAVY_CAP + 1), s);
    }
    GEM_HIP_CHECK(hipHostMalloc((void **)&h_hdr, hdr_bytes, hipHostMallocDefault));
    for (int i = 0; i < 3; ++i) {
      items[i].alloc(sizeof(VItem) * cap, s);
      blk[i].alloc(8ull * (cap + 1), s);
      GEM_HIP_CHECK(hipHostMalloc((void **)&h_items[i], sizeof(VItem) * cap, hipHostMallocDefault));
      GEM_HIP_CHECK(hipHostMalloc((void **)&h_blk[i], 8ull * (cap + 1), hipHostMallocDefault));
      std::memset(h_items[i], 0, sizeof(VItem) * cap);
      std::memset(h_blk[i], 0, 8ull * (cap + 1));
    }
    std::memset(h_hdr, 0, hdr_bytes);
  }
  ~ExtSlot() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_hdr) (void)hipHostFree(h_hdr);
    for (int i = 0; i < 3; ++i) {
      if (h_items[i]) (void)hipHostFree(h_items[i]);
      if (h_blk[i]) (void)hipHostFree(h_blk[i]);
    }
  }
  uint32_t *nitems(int i) { return reinterpret_cast<uint32_t *>(h_hdr + HDR_NITEMS) + i; }
  uint64_t *counters() { return reinterpret_cast<uint64_t *>(h_hdr + HDR_CTR); }
  uint64_t *pitems() { return reinterpret_cast<uint64_t *>(h_hdr + HDR_PITEMS); }
  unsigned long long *dev_counter(size_t i) {
    return reinterpret_cast<unsigned long long *>(hdr.as<unsigned char>() + HDR_CTR) + i;
  }

  // The step's work, enqueued on the stream (eagerly or under capture).
  void enqueue(gem_ctx *c, bool prof_kernels, bool count_items) {
    hipStream_t s = c->stream;
    unsigned char *d = hdr.as<unsigned char>();
    GEM_HIP_CHECK(hipMemcpyAsync(d, h_hdr, hdr_bytes, hipMemcpyHostToDevice, s));
    for (int i = 0; i < 3; ++i) {
      GEM_HIP_CHECK(hipMemcpyAsync(items[i].p, h_items[i], sizeof(VItem) * cap, hipMemcpyHostToDevice, s));
      GEM_HIP_CHECK(hipMemcpyAsync(blk[i].p, h_blk[i], 8ull * (cap + 1), hipMemcpyHostToDevice, s));
    }
    const uint32_t *dn = reinterpret_cast<const uint32_t *>(d + HDR_NITEMS);
    unsigned int *dq = reinterpret_cast<unsigned int *>(d + HDR_QUEUE);
    unsigned int *hn = reinterpret_cast<unsigned int *>(d + HDR_HEAVYN);
    launch_kernels<false>(c, grid, dn, dq, dq + 3, hn, prof_kernels, count_items, 0, ~0ull, 0, ~0ull, 0, ~0ull);
    GEM_HIP_CHECK(hipMemcpyAsync(h_hdr, d, hdr_bytes, hipMemcpyDeviceToHost, s));
  }

  // The three kernels over task ranges [b_i, e_i) of launch i (queues
  // dq[0..2]), then the heavy rows the two walking launches set aside (queues
  // hq[0..1], row counts hn[0..1], all starting at zero).
  template <bool OOC>
  void launch_kernels(gem_ctx *c, uint32_t grd, const uint32_t *dn, unsigned int *dq, unsigned int *hq,
                      unsigned int *hn, bool prof_kernels, bool count_items, uint64_t b0, uint64_t e0, uint64_t b1,
                      uint64_t e1, uint64_t b2, uint64_t e2) {
    hipStream_t s = c->stream;
    unsigned long long *dp = reinterpret_cast<unsigned long long *>(hdr.as<unsigned char>() + HDR_PITEMS);
    if (b0 < e0) {
      ProfLaunch pl(c, prof_kernels ? K_VEXT_WRITE : -1);
      k_vext<false, OOC><<<grd * (VB / vthreads), vthreads, lds, s>>>(items[0].as<VItem>(), dn + 0, blk[0].as<uint64_t>(), dq + 0, max_nb,
                                               max_ndup, count_items ? dp : nullptr, count_items ? dp + 4 : nullptr,
                                               b0, e0, heavy[0].as<HeavyRow>(), hn + 0, HEAVY_CAP, heavy_t);
      GEM_LAUNCH_CHECK("k_vext<write>");
    }
    if (b1 < e1) {
      ProfLaunch pl(c, prof_kernels ? K_VEXT_FINAL : -1);
      k_vext<true, OOC><<<grd * (VB / vthreads), vthreads, lds, s>>>(items[1].as<VItem>(), dn + 1, blk[1].as<uint64_t>(), dq + 1, max_nb,
                                              max_ndup, count_items ? dp + 1 : nullptr, nullptr, b1, e1,
                                              heavy[1].as<HeavyRow>(), hn + 1, HEAVY_CAP, heavy_t);
      GEM_LAUNCH_CHECK("k_vext<final>");
    }
    if (b2 < e2) {
      ProfLaunch pl(c, prof_kernels ? K_VCOUNT1 : -1);
      k_vcount1<OOC><<<grd, VB, 0, s>>>(items[2].as<VItem>(), dn + 2, blk[2].as<uint64_t>(), dq + 2, b2, e2);
      GEM_LAUNCH_CHECK("k_vcount1");
    }
    if (b0 < e0) {
      ProfLaunch pl(c, prof_kernels ? K_VEXT_HEAVY_WRITE : -1);
      k_heavy_scan<<<1, 1024, 0, s>>>(heavy[0].as<HeavyRow>(), hn + 0, HEAVY_CAP, heavy_chunk, hpre[0].as<uint64_t>());
      GEM_LAUNCH_CHECK("k_heavy_scan");
      k_vext_heavy<false, OOC><<<grd, VB, 0, s>>>(items[0].as<VItem>(), heavy[0].as<HeavyRow>(), hn + 0, HEAVY_CAP,
                                                   hpre[0].as<uint64_t>(), hq + 0, heavy_chunk,
                                                   count_items ? dp + 2 : nullptr, count_items ? dp + 5 : nullptr);
      GEM_LAUNCH_CHECK("k_vext_heavy<write>");
    }
    if (b1 < e1) {
      ProfLaunch pl(c, prof_kernels ? K_VEXT_HEAVY_FINAL : -1);
      k_heavy_scan<<<1, 1024, 0, s>>>(heavy[1].as<HeavyRow>(), hn + 1, HEAVY_CAP, heavy_chunk, hpre[1].as<uint64_t>());
      GEM_LAUNCH_CHECK("k_heavy_scan");
      k_vext_heavy<true, OOC><<<grd, VB, 0, s>>>(items[1].as<VItem>(), heavy[1].as<HeavyRow>(), hn + 1, HEAVY_CAP,
                                                  hpre[1].as<uint64_t>(), hq + 1, heavy_chunk,
                                                  count_items ? dp + 3 : nullptr, nullptr);
      GEM_LAUNCH_CHECK("k_vext_heavy<final>");
    }
  }

  // Streamed tables: the same step, its tasks cut into groups whose table
  // blocks are copied in while the previous group runs (eager launches only).
  void launch_streamed(gem_ctx *c, TableCache &tc) {
    hipStream_t s = c->stream;
    unsigned char *d = hdr.as<unsigned char>();
    GEM_HIP_CHECK(hipMemcpyAsync(d, h_hdr, hdr_bytes, hipMemcpyHostToDevice, s));
    uint64_t T[3], base[3];
    uint64_t total = 0;
    for (int i = 0; i < 3; ++i) {
      GEM_HIP_CHECK(hipMemcpyAsync(items[i].p, h_items[i], sizeof(VItem) * cap, hipMemcpyHostToDevice, s));
      GEM_HIP_CHECK(hipMemcpyAsync(blk[i].p, h_blk[i], 8ull * (cap + 1), hipMemcpyHostToDevice, s));
      T[i] = *nitems(i) ? h_blk[i][*nitems(i)] : 0;
      base[i] = total;
      total += T[i];
    }
    if (total >= 0xFFFFFFFFull) throw Error{GEM_E_LIMIT, "too many tasks in one streamed iteration"};
    // The iteration is planned in windows of tasks: a block's marks span its
    // first to last use inside the window, so a window whose blocks cannot
    // all be placed is halved (down to one task) and planned again.
    const uint32_t *dn = reinterpret_cast<const uint32_t *>(d + HDR_NITEMS);
    const bool prof = c->prof.on;
    uint64_t wa = 0, W = total;
    while (wa < total) {
      const uint64_t wb = std::min(total, wa + W);
      tc.reset_marks();
      for (int i = 0; i < 3; ++i) {
        if (!T[i] || base[i] + T[i] <= wa || base[i] >= wb) continue;
        const uint64_t lo = std::max(wa, base[i]), hi = std::min(wb, base[i] + T[i]);
        k_vmark<<<(uint32_t)std::min<uint64_t>(hi - lo, 4096), VB, 0, s>>>(
            items[i].as<VItem>(), *nitems(i), blk[i].as<uint64_t>(), base[i], wa, wb, i == 2, tc.d_first(),
            tc.d_last());
        GEM_LAUNCH_CHECK("k_vmark");
      }
      if (!tc.plan(wb - wa)) {
        if (W == 1)
          throw Error{GEM_E_LIMIT, "the table blocks read by one 256-row task do not fit the device cache"};
        W = std::max<uint64_t>(1, W / 2);
        continue;
      }
      const auto &G = tc.groups();
      unsigned int *q = tc.queues();
      for (size_t g = 0; g < G.size(); ++g) {
        tc.issue(g);
        uint64_t lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
          lo[i] = std::max(wa + G[g].a, base[i]);
          hi[i] = std::min(wa + G[g].b, base[i] + T[i]);
          if (lo[i] < hi[i]) {
            lo[i] -= base[i];
            hi[i] -= base[i];
          } else {
            lo[i] = hi[i] = 0;
          }
        }
        const uint64_t most = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        const uint32_t grd = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(grid, most));
        // per group: 3 main queues, 2 heavy queues, 2 heavy-row counts
        unsigned int *gq = q + 7 * g;
        launch_kernels<true>(c, grd, dn, gq, gq + 3, gq + 5, prof, prof, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
        tc.done(g);
      }
      wa = wb;
      W = std::min(total, 2 * W);
    }
    GEM_HIP_CHECK(hipMemcpyAsync(h_hdr, d, hdr_bytes, hipMemcpyDeviceToHost, s));
  }

  void launch(gem_ctx *c, bool use_graph) {
    if (!use_graph) {
      enqueue(c, c->prof.on, c->prof.on);
      return;
    }
    if (!exec) {
      hipStream_t s = c->stream;
      GEM_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(c, false, false);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      GEM_HIP_CHECK(hipStreamEndCapture(s, &graph));
      GEM_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    GEM_HIP_CHECK(hipGraphLaunch(exec, c->stream));
  }
};

}  // namespace

namespace {

// Device side of the scheduler (host/vsched.h): level buffers from the
// caching allocator, length passes, extend steps through the fixed launch
// buffers of the run (eager, captured as a hipGraph, or with streamed tables).
struct HipSchedBackend {
  using Buf = DevBuf;
  using Item = gemh::SItem<DevBuf>;
  gem_ctx *c;
  std::vector<VQueryState *> &qs;
  ExtSlot &slot;
  TableCache *tc;
  bool verbose;
  uint64_t cap_;
  LazyStates *lazy = nullptr;
  using clk = std::chrono::steady_clock;
  clk::time_point last_print = clk::now();
  double t_len = 0, t_ext = 0;
  uint64_t ext_tasks = 0;

  static uint64_t buf_bytes(const DevBuf &b) { return b.bytes; }

  std::shared_ptr<DevBuf> alloc(uint64_t bytes) {
    auto b = std::make_shared<DevBuf>();
    b->alloc(bytes, c->stream);
    return b;
  }

  std::shared_ptr<DevBuf> shrink(const std::shared_ptr<DevBuf> &from, uint64_t bytes) {
    auto b = std::make_shared<DevBuf>();
    b->alloc(bytes, c->stream);
    GEM_HIP_CHECK(hipMemcpyAsync(b->p, from->p, bytes, hipMemcpyDeviceToDevice, c->stream));
    return b;
  }

  VItem vitem(const Item &x) const {
    VItem it{};
    it.lv = qs[x.q]->levels.as<VLevel>() + x.level;
    it.rows = x.rows->as<uint32_t>();
    it.row_begin = x.row_begin;
    it.row_end = x.row_end;
    it.k = x.k;
    return it;
  }

  void length_pass(const std::vector<Item> &items, std::vector<std::vector<uint64_t>> *prefix) {
    const auto t0 = clk::now();
    hipStream_t s = c->stream;
    Launch L;
    for (const Item &x : items) L.add(vitem(x), qs[x.q]->max_nb[x.level], qs[x.q]->max_ndup[x.level]);
    DevLaunch D;
    upload(c, L, D);
    const uint64_t nblk = L.blk.back(), nunit = nblk * VW;
    DevBuf bu, bp;
    bu.alloc(nunit * 8, s);
    bp.alloc((nunit + 1) * 8, s);
    {
      ProfLaunch pl(c, K_VLEN);
      k_vlen<<<grid_for(nblk), VB, 0, s>>>(D.items.as<VItem>(), (uint32_t)L.items.size(), D.blk.as<uint64_t>(),
                                            D.queue.as<unsigned int>(), bu.as<uint64_t>());
      GEM_LAUNCH_CHECK("k_vlen");
    }
    scan_u64(c, bu.as<uint64_t>(), bp.as<uint64_t>(), nunit);
    uint64_t *pref = pinned<uint64_t>(c, nunit + 1);
    GEM_HIP_CHECK(hipMemcpyAsync(pref, bp.p, (nunit + 1) * 8, hipMemcpyDeviceToHost, s));
    GEM_HIP_CHECK(hipStreamSynchronize(s));
    prefix->resize(items.size());
    for (size_t t = 0; t < items.size(); ++t) {
      const Item &x = items[t];
      const uint64_t u0 = L.blk[t] * VW;
      const uint64_t nu = (x.row_end - x.row_begin + gemh::SCHED_UNIT - 1) / gemh::SCHED_UNIT;
      std::vector<uint64_t> &P = (*prefix)[t];
      P.resize(nu + 1);
      for (uint64_t i = 0; i <= nu; ++i) P[i] = pref[u0 + i] - pref[u0];
      // rows + one offset pair per backward list
      prof_bytes(c, K_VLEN, (double)(x.row_end - x.row_begin) * (4.0 * x.k + 16.0 * qs[x.q]->max_nb[x.level]));
    }
    t_len += std::chrono::duration<double>(clk::now() - t0).count();
  }

  void extend(const std::vector<Item> &wr_items, const std::vector<Item> &fin_items, std::vector<uint64_t> *got_wr,
              std::vector<uint64_t> *got_fin) {
    const auto t0 = clk::now();
    hipStream_t s = c->stream;
    // launch tables: write walks, final walks, final single-list counts; one
    // u64 result counter per item, in the device header
    Launch wr, fin, fin1;
    std::vector<std::pair<int, size_t>> fin_at(fin_items.size());
    size_t nctr = 0;
    for (const Item &x : wr_items) {
      VItem it = vitem(x);
      it.out = x.out->as<uint32_t>();
      it.out_cap = x.out_rows;
      it.counter = slot.dev_counter(nctr++);
      wr.add(it, qs[x.q]->max_nb[x.level], qs[x.q]->max_ndup[x.level]);
    }
    std::vector<size_t> fin_ctr(fin_items.size());
    for (size_t t = 0; t < fin_items.size(); ++t) {
      const Item &x = fin_items[t];
      VItem it = vitem(x);
      fin_ctr[t] = nctr;
      it.counter = slot.dev_counter(nctr++);
      if (qs[x.q]->max_nb[x.level] == 1)
        fin1.add(it, 1, qs[x.q]->max_ndup[x.level]);
      else
        fin.add(it, qs[x.q]->max_nb[x.level], qs[x.q]->max_ndup[x.level]);
    }
    if (nctr > slot.cap) throw Error{GEM_E_STATE, "more vjoin tasks than launch slots"};
    Launch *L3[3] = {&wr, &fin, &fin1};
    for (int i = 0; i < 3; ++i) {
      Launch &L = *L3[i];
      if (L.items.size() > slot.cap) throw Error{GEM_E_STATE, "more vjoin items than launch slots"};
      for (size_t t = 0; t < L.items.size(); ++t) slot.h_items[i][t] = L.items[t];
      for (size_t t = 0; t < L.blk.size(); ++t) slot.h_blk[i][t] = L.blk[t];
      if (L.blk.empty()) slot.h_blk[i][0] = 0;
      *slot.nitems(i) = (uint32_t)L.items.size();
      ext_tasks += L.blk.empty() ? 0 : L.blk.back();
    }
    // counters / work queues / walked-entry counts travel zeroed in the copy
    std::memset(slot.h_hdr + ExtSlot::HDR_QUEUE, 0, slot.hdr_bytes - ExtSlot::HDR_QUEUE);
    const bool prof = c->prof.on;
    if (tc)
      slot.launch_streamed(c, *tc);
    else
      slot.launch(c, c->use_graph && !prof);
    GEM_HIP_CHECK(hipStreamSynchronize(s));
    if (tc) tc->finish();
    const uint64_t *hc = slot.counters();
    got_wr->resize(wr_items.size());
    got_fin->resize(fin_items.size());
    for (size_t t = 0; t < wr_items.size(); ++t) {
      (*got_wr)[t] = hc[t];
      if (hc[t] > wr_items[t].out_rows) throw Error{GEM_E_STATE, "vjoin output exceeded its bound"};
    }
    for (size_t t = 0; t < fin_items.size(); ++t) (*got_fin)[t] = hc[fin_ctr[t]];
    if (prof) {
      const uint64_t *pc = slot.pitems();
      // algorithmic bytes: rows read (4 B per column + an offset pair per
      // backward list), 4 B per list entry walked, output rows written
      auto rows_bytes = [&](const Launch &L) {
        double b = 0;
        for (const auto &it : L.items)
          b += (double)(it.row_end - it.row_begin) * (4.0 * it.k + 16.0 * (double)L.max_nb);
        return b;
      };
      // (the heavy-row kernels: the entries they walk and the rows they write)
      prof_bytes(c, K_VEXT_WRITE, rows_bytes(wr) + 4.0 * (double)pc[0] + (double)pc[4]);
      prof_bytes(c, K_VEXT_FINAL, rows_bytes(fin) + 4.0 * (double)pc[1]);
      prof_bytes(c, K_VCOUNT1, rows_bytes(fin1));
      prof_bytes(c, K_VEXT_HEAVY_WRITE, 4.0 * (double)pc[2] + (double)pc[5]);
      prof_bytes(c, K_VEXT_HEAVY_FINAL, 4.0 * (double)pc[3]);
    }
    prof_flush(c);
    t_ext += std::chrono::duration<double>(clk::now() - t0).count();
  }

  // Swapping: a level's rows to pinned host memory and back, both copies on
  // the context stream (the device buffer goes back to the pool at once: its
  // next user is ordered after the copy).  A host buffer is freed after the
  // stream has drained, so no copy still reads or writes it.
  std::shared_ptr<void> swap_out(const DevBuf &b, uint64_t bytes) {
    void *h = nullptr;
    GEM_HIP_CHECK(hipHostMalloc(&h, std::max<uint64_t>(bytes, 1), hipHostMallocDefault));
    hipStream_t s = c->stream;
    std::shared_ptr<void> hp(h, [s](void *p) {
      (void)hipStreamSynchronize(s);
      (void)hipHostFree(p);
    });
    GEM_HIP_CHECK(hipMemcpyAsync(h, b.p, bytes, hipMemcpyDeviceToHost, s));
    return hp;
  }
  std::shared_ptr<DevBuf> swap_in(const std::shared_ptr<void> &h, uint64_t bytes) {
    auto b = std::make_shared<DevBuf>();
    b->alloc(bytes, c->stream);
    GEM_HIP_CHECK(hipMemcpyAsync(b->p, h.get(), bytes, hipMemcpyHostToDevice, c->stream));
    return b;
  }

  bool ready(uint32_t qi) {
    const VQueryState &V = *qs[qi];
    return !V.pending || V.twin || !lazy || lazy->ready(qi);
  }
  void wait_ready(uint32_t qi) {
    const VQueryState &V = *qs[qi];
    if (V.pending && !V.twin && lazy) lazy->complete(qi);
  }

  // A pending state is completed when it starts: a primary by its preparer
  // (ranked orders, plan), a twin from its partner (which started before it).
  // Returns false when there is nothing to run (a twin of a query with one
  // candidate order).
  bool complete(uint32_t qi) {
    VQueryState &V = *qs[qi];
    if (!V.pending) return true;
    if (V.twin) {
      const VQueryState &P = *qs[(size_t)V.s.partner];
      if (P.pending || P.ranked.size() < 2) return false;
      V.first_attempt = P.first_attempt;
      V.growth = P.growth;
      V.width = P.width;
      V.max_attempts = P.max_attempts;
      V.ranked = P.ranked;
      V.ranked_idx = P.ranked_idx;
      V.mnb_all = P.mnb_all;
      V.mnd_all = P.mnd_all;
    } else {
      if (!lazy) throw Error{GEM_E_STATE, "a pending query without a preparer"};
      lazy->complete(qi);
    }
    if (V.mnb_all > slot.max_nb || V.mnd_all > slot.max_ndup)
      throw Error{GEM_E_STATE, "a query needs more launch LDS than the run was sized for"};
    V.pending = false;
    return true;
  }

  void admit(uint32_t qi) {
    VQueryState &V = *qs[qi];
    if (!complete(qi)) {
      V.s.done = true;
      V.s.cancelled = true;
      return;
    }
    if (!V.twin) {
      // a primary completed now: its levels' list counts and first limit
      V.s.nb = V.max_nb;
      if (!V.s.attempts) V.s.attempt_limit = attempt_limit(V, 0);
    }
    if (V.twin && V.s.partner >= 0) {
      // a racing twin starts at its partner's current limit, on the order
      // after the one its partner runs
      const VQueryState &P = *qs[(size_t)V.s.partner];
      const uint32_t K = (uint32_t)V.ranked.size();
      const uint32_t a = P.s.attempts;
      V.shift = (P.cur + 1 + K - a % K) % K;
      V.cur = (a + V.shift) % K;
      replan(c, V, V.ranked[V.cur]);
      V.s.nb = V.max_nb;
      if (V.mnb_all > slot.max_nb || V.mnd_all > slot.max_ndup)
        throw Error{GEM_E_STATE, "racing order needs more launch LDS than the run was sized for"};
      V.s.attempts = a;
      V.s.attempt_start = V.s.work;
      V.s.attempt_limit = attempt_limit(V, a);
      if (verbose) {
        std::fprintf(stderr, "[vjoin] query %u races query %lld: order %u, limit %.3g\n", qi,
                     (long long)V.s.partner, V.ranked_idx[V.cur], (double)V.s.attempt_limit);
        std::fflush(stderr);
      }
    }
    seed_query(c, V);
  }

  // The attempt of query qi went past its limit with its levels still
  // pending.  It goes on when it has counted at least a quarter of the cap and,
  // at the rate it counted so far, needs less work to pass the cap than 4/5 of
  // the next attempt's limit (a new attempt starts from nothing, and an
  // attempt whose cost is near its limit would otherwise end or restart by
  // chance: the order rows come out of a level in varies from run to run).
  // Its limit becomes what it has spent plus 1.5 times that estimate.
  bool prolong(uint32_t qi) {
    VQueryState &V = *qs[qi];
    gemh::SQuery<DevBuf> &Q = V.s;
    if (!cap_ || Q.count < cap_ / 4 || Q.count == 0) return false;
    const uint64_t spent = Q.work - Q.attempt_start;
    const double need = (double)spent * ((double)(cap_ + 1) / (double)Q.count - 1.0);
    const uint64_t next = attempt_limit(V, Q.attempts + 1);
    if (next && need >= 0.8 * (double)next) return false;
    Q.attempt_limit = spent + (uint64_t)(1.5 * need) + 1;
    if (verbose) {
      std::fprintf(stderr, "[vjoin] query %u: attempt %u goes on (%.0f%% of the cap counted), limit %.3g\n", qi,
                   Q.attempts, 100.0 * (double)Q.count / (double)cap_, (double)Q.attempt_limit);
      std::fflush(stderr);
    }
    return true;
  }

  // The attempt of query qi went past its limit (its levels are dropped):
  // the next ranked order, with the next attempt's limit.
  bool restart(uint32_t qi) {
    VQueryState &V = *qs[qi];
    gemh::SQuery<DevBuf> &Q = V.s;
    if (!attempt_limit(V, 0)) return false;
    const uint32_t a = Q.attempts + 1;
    if (V.max_attempts && a >= V.max_attempts) return false;
    // the ended attempt's progress: embeddings counted per unit of work
    const uint32_t K = (uint32_t)V.ranked.size();
    if (V.rate.size() != K) V.rate.assign(K, 0.0);
    const uint64_t spent = Q.work - Q.attempt_start;
    if (spent) V.rate[V.cur] = std::max(V.rate[V.cur], (double)Q.count / (double)spent);
    // every ranked order once, then the one that counted fastest (the ranking
    // again while none has counted anything)
    V.cur = (a + V.shift) % K;
    if (a >= K) {
      uint32_t best = 0;
      for (uint32_t j = 1; j < K; ++j)
        if (V.rate[j] > V.rate[best]) best = j;
      if (V.rate[best] > 0) V.cur = best;
    }
    replan(c, V, V.ranked[V.cur]);
    if (tc) tc->bind(V, qi);
    if (V.mnb_all > slot.max_nb || V.mnd_all > slot.max_ndup)
      throw Error{GEM_E_STATE, "restart order needs more launch LDS than the run was sized for"};
    Q.attempts = a;
    Q.attempt_start = Q.work;
    Q.attempt_limit = attempt_limit(V, a);
    seed_query(c, V);
    if (verbose) {
      std::fprintf(stderr, "[vjoin] query %u: attempt %u with order %u, limit %.3g\n", qi, a, V.ranked_idx[V.cur],
                   (double)Q.attempt_limit);
      std::fflush(stderr);
    }
    return true;
  }

  // wall time and iterations by the number of queries active (tail: at most 4)
  clk::time_point last_iter = clk::now();
  size_t last_active = 0;
  double t_tail = 0, t_body = 0;
  uint64_t it_tail = 0, it_body = 0;

  void on_iteration(uint64_t iter, size_t active, uint64_t live, uint64_t budget) {
    const auto now = clk::now();
    if (last_active) {
      const double dt = std::chrono::duration<double>(now - last_iter).count();
      if (last_active <= 4) {
        t_tail += dt;
        ++it_tail;
      } else {
        t_body += dt;
        ++it_body;
      }
    }
    last_iter = now;
    last_active = active;
    if (!verbose || now - last_print < std::chrono::seconds(2)) return;
    last_print = clk::now();
    std::fprintf(stderr, "[vjoin] iteration %llu: %zu of %zu queries active, pending levels %.2f of %.2f GiB\n",
                 (unsigned long long)iter, active, qs.size(), live / 1073741824.0, budget / 1073741824.0);
    std::fflush(stderr);
  }
};

}  // namespace

void VBatchRunner::run(std::vector<VQueryState *> &qs) {
  g_persistent_blocks = c_->persistent_blocks;
  // the level budget: given, or a quarter of what the device could still
  // give the run's level buffers (free memory plus empty pool segments)
  uint64_t total_budget = budget_bytes_;
  if (!total_budget) {
    size_t free_b = 0, total_b = 0;
    GEM_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    total_budget = std::max<uint64_t>((free_b + pool_reusable_bytes()) / 4, 1ull << 30);
  }
  total_budget = std::max<uint64_t>(total_budget, 1ull << 20);
  // fixed launch buffers for the run (one item per query per launch at most)
  ExtSlot slot;
  {
    uint32_t mnb = 1, mnd = 0;
    for (VQueryState *S : qs) {
      for (size_t i = 0; i < S->max_nb.size(); ++i) {
        mnb = std::max(mnb, S->max_nb[i]);
        mnd = std::max(mnd, S->max_ndup[i]);
      }
      // every order a restart may switch to
      if (attempt_limit(*S, 0)) {
        mnb = std::max(mnb, S->mnb_all);
        mnd = std::max(mnd, S->mnd_all);
      }
    }
    slot.init(c_, (uint32_t)qs.size(), mnb, mnd);
  }
  // streamed tables: a block cache shared by the run's queries
  std::unique_ptr<TableCache> tc;
  if (streamed_) {
    if (!c_->stream_cache_bytes) throw Error{GEM_E_STATE, "streamed join without a table cache size"};
    tc = std::make_unique<TableCache>(c_, c_->stream_cache_bytes, c_->stream_slot_bytes);
    tc->attach(qs);
  }
  // queries are seeded with the identity level of their first vertex when
  // the scheduler starts them (HipSchedBackend::admit)
  std::vector<gemh::SQuery<DevBuf> *> sq;
  for (VQueryState *S : qs) {
    gemh::SQuery<DevBuf> &Q = S->s;
    const int64_t partner = Q.partner;
    Q = gemh::SQuery<DevBuf>{};
    Q.partner = partner;
    // (a pending state's query may still be being prepared: its vertex count
    // was recorded when the state was made)
    Q.nv = S->pending ? S->pending_nv : S->qd->Q.nv;
    Q.nb = S->pending ? std::vector<uint32_t>(Q.nv, S->max_nb.empty() ? 1u : S->max_nb[0]) : S->max_nb;
    Q.level_rows.assign(Q.nv, 0);
    Q.attempt_limit = S->pending ? 0 : attempt_limit(*S, 0);
    sq.push_back(&Q);
  }
  const char *verbose_env = std::getenv("GEM_VERBOSE");
  const bool verbose = verbose_env && verbose_env[0] == '1';
  HipSchedBackend be{c_, qs, slot, tc.get(), verbose, cap_, lazy_};
  gemh::SchedConfig cfg;
  cfg.cap = cap_;
  cfg.work_limit = work_limit_;
  cfg.budget = total_budget;
  cfg.max_active = c_->max_active;
  cfg.primaries = primaries_;
  cfg.swap_bytes = c_->swap_bytes;
  const auto t0 = std::chrono::steady_clock::now();
  stats_ = gemh::SchedStats{};
  gemh::run_schedule(be, sq, cfg, &stats_);
  c_->last_sched = stats_;
  c_->last_budget = total_budget;
  if (verbose) {
    std::fprintf(stderr,
                 "[vjoin] run of %zu queries: %llu iterations in %.2f s: %llu length passes %.2f s, extend steps "
                 "%.2f s (%llu 256-row tasks); budget %.2f GiB, peak charged %.2f GiB, %llu forced iterations, "
                 "%llu restarts, %llu queries started (at most %u at once); %llu iterations with more than 4 "
                 "queries active %.2f s, %llu with at most 4 %.2f s\n",
                 qs.size(), (unsigned long long)stats_.iterations,
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                 (unsigned long long)stats_.length_passes, be.t_len, be.t_ext, (unsigned long long)be.ext_tasks,
                 total_budget / 1073741824.0, stats_.peak_charged / 1073741824.0,
                 (unsigned long long)stats_.forced, (unsigned long long)stats_.restarts,
                 (unsigned long long)stats_.admitted, cfg.max_active, (unsigned long long)be.it_body, be.t_body,
                 (unsigned long long)be.it_tail, be.t_tail);
    std::fflush(stderr);
  }
}

}  // namespace gemd
