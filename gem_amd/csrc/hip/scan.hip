// Exclusive prefix sums used by the filter (candidate ranks), the edge tables
// (row offsets) and the join (output offsets).  Reduce-then-scan: per-tile sums,
// a one-block scan of the tile sums, then a tile-local scan with the carried
// base.  Integer adds only, so results are exact and order-independent.
#include "ctx.h"
#include "wave.h"

namespace gemd {

constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_ITEMS = 16;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums_u64(const uint64_t *in, uint64_t n, uint64_t *sums) {
  __shared__ uint64_t sm[8];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint64_t acc = 0;
  for (uint32_t i = 0; i < SCAN_ITEMS; ++i) {
    uint64_t idx = base + (uint64_t)i * SCAN_THREADS + threadIdx.x;
    if (idx < n) acc += in[idx];
  }
  uint64_t ex;
  uint64_t tot = block_excl_scan_u64(acc, &ex, sm);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// one block: exclusive scan of m values, out[m] = total
__global__ void __launch_bounds__(1024) k_scan_small_u64(const uint64_t *in, uint64_t m, uint64_t *out) {
  __shared__ uint64_t sm[16];
  uint64_t carry = 0;
  for (uint64_t b = 0; b < m; b += blockDim.x) {
    uint64_t idx = b + threadIdx.x;
    uint64_t x = idx < m ? in[idx] : 0;
    uint64_t ex;
    uint64_t tot = block_excl_scan_u64(x, &ex, sm);
    if (idx < m) out[idx] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) out[m] = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_scan_u64(const uint64_t *in, uint64_t n, const uint64_t *tile_base,
                                                              uint64_t *out) {
  __shared__ uint64_t sm[8];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  // each thread owns SCAN_ITEMS consecutive elements
  uint64_t v[SCAN_ITEMS];
  uint64_t acc = 0;
  const uint64_t my = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
  for (uint32_t i = 0; i < SCAN_ITEMS; ++i) {
    uint64_t idx = my + i;
    v[i] = idx < n ? in[idx] : 0;
    acc += v[i];
  }
  uint64_t ex;
  block_excl_scan_u64(acc, &ex, sm);
  uint64_t run = tile_base[blockIdx.x] + ex;
#pragma unroll
  for (uint32_t i = 0; i < SCAN_ITEMS; ++i) {
    uint64_t idx = my + i;
    if (idx < n) out[idx] = run;
    run += v[i];
  }
}

void scan_u64(gem_ctx *c, const uint64_t *d_in, uint64_t *d_out, uint64_t n) {
  if (n <= 4 * 1024 * 16) {
    k_scan_small_u64<<<1, 1024, 0, c->stream>>>(d_in, n, d_out);
    GEM_LAUNCH_CHECK("k_scan_small_u64");
    return;
  }
  const uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  DevBuf sums, base;
  sums.alloc(tiles * 8, c->stream);
  base.alloc((tiles + 1) * 8, c->stream);
  k_tile_sums_u64<<<(uint32_t)tiles, SCAN_THREADS, 0, c->stream>>>(d_in, n, sums.as<uint64_t>());
  GEM_LAUNCH_CHECK("k_tile_sums_u64");
  if (tiles > 4 * 1024 * 16) {
    scan_u64(c, sums.as<uint64_t>(), base.as<uint64_t>(), tiles);
  } else {
    k_scan_small_u64<<<1, 1024, 0, c->stream>>>(sums.as<uint64_t>(), tiles, base.as<uint64_t>());
    GEM_LAUNCH_CHECK("k_scan_small_u64");
  }
  k_tile_scan_u64<<<(uint32_t)tiles, SCAN_THREADS, 0, c->stream>>>(d_in, n, base.as<uint64_t>(), d_out);
  GEM_LAUNCH_CHECK("k_tile_scan_u64");
  // out[n] = total
  GEM_HIP_CHECK(hipMemcpyAsync(d_out + n, base.as<uint64_t>() + tiles, 8, hipMemcpyDeviceToDevice, c->stream));
}

// Many independent in-place scans in one launch: block b scans
// segs[b].p[0..n) (u64, exclusive), stores the total at segs[b].p[n] and in
// totals[b].  Segments longer than the library's one-block limit are left to
// scan_u64 (their total is gathered by k_gather_totals).
__global__ void __launch_bounds__(1024) k_scan_many_u64(const ScanSeg *__restrict__ segs, uint64_t small_max,
                                                       uint64_t *__restrict__ totals) {
  __shared__ uint64_t sm[16];
  const ScanSeg sg = segs[blockIdx.x];
  if (sg.n > small_max) return;
  uint64_t carry = 0;
  for (uint64_t b = 0; b < sg.n; b += blockDim.x) {
    const uint64_t idx = b + threadIdx.x;
    const uint64_t x = idx < sg.n ? sg.p[idx] : 0;
    uint64_t ex;
    const uint64_t tot = block_excl_scan_u64(x, &ex, sm);
    if (idx < sg.n) sg.p[idx] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    sg.p[sg.n] = carry;
    totals[blockIdx.x] = carry;
  }
}

__global__ void k_gather_totals(const ScanSeg *__restrict__ segs, uint32_t nseg, uint64_t small_max,
                                uint64_t *__restrict__ totals) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nseg && segs[b].n > small_max) totals[b] = segs[b].p[segs[b].n];
}

void scan_many_u64(gem_ctx *c, const ScanSeg *d_segs, const ScanSeg *h_segs, uint32_t nseg, uint64_t *d_totals) {
  if (!nseg) return;
  constexpr uint64_t SMALL = 4 * 1024 * 16;  // scan_u64's one-block limit
  k_scan_many_u64<<<nseg, 1024, 0, c->stream>>>(d_segs, SMALL, d_totals);
  GEM_LAUNCH_CHECK("k_scan_many_u64");
  bool big = false;
  for (uint32_t b = 0; b < nseg; ++b)
    if (h_segs[b].n > SMALL) {
      scan_u64(c, h_segs[b].p, h_segs[b].p, h_segs[b].n);
      big = true;
    }
  if (big) {
    k_gather_totals<<<(nseg + 255) / 256, 256, 0, c->stream>>>(d_segs, nseg, SMALL, d_totals);
    GEM_LAUNCH_CHECK("k_gather_totals");
  }
}

// One block per segment: exclusive scan of len u32 counts, out[len] = total.
__global__ void __launch_bounds__(1024) k_scan_segments_u32(const uint32_t *in, uint32_t *out, uint64_t len) {
  __shared__ uint64_t sm[16];
  const uint32_t *src = in + (uint64_t)blockIdx.x * len;
  uint32_t *dst = out + (uint64_t)blockIdx.x * (len + 1);
  uint64_t carry = 0;
  for (uint64_t b = 0; b < len; b += blockDim.x) {
    uint64_t idx = b + threadIdx.x;
    uint64_t x = idx < len ? src[idx] : 0;
    uint64_t ex;
    uint64_t tot = block_excl_scan_u64(x, &ex, sm);
    if (idx < len) dst[idx] = (uint32_t)(carry + ex);
    carry += tot;
  }
  if (threadIdx.x == 0) dst[len] = (uint32_t)carry;
}

void scan_segments_u32(gem_ctx *c, const uint32_t *d_in, uint32_t *d_out, uint32_t nseg, uint64_t len) {
  if (nseg == 0) return;
  k_scan_segments_u32<<<nseg, 1024, 0, c->stream>>>(d_in, d_out, len);
  GEM_LAUNCH_CHECK("k_scan_segments_u32");
}

}  // namespace gemd
