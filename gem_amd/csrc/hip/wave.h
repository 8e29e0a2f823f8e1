// Wave64 building blocks for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gemd {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Load / store through the global address space.  Pointers read from memory
// (list addresses, row pointers) are generic to the compiler, which then
// emits flat instructions: those also count against the LDS wait counter, so
// every wait on them waits for LDS traffic too.  All such pointers here are
// device memory or mapped host memory, which global instructions reach.
template <class T>
__device__ __forceinline__ T ldg(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ void stg(T *p, T v) {
  *(__attribute__((address_space(1))) T *)p = v;
}

// popcount of the bits of `m` below this lane (v_mbcnt_lo/hi)
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t x, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)lane);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, uint32_t lane) {
  uint32_t lo = readlane_u32((uint32_t)x, lane), hi = readlane_u32((uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive wave scan (Hillis-Steele over 64 lanes).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (l >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if (l >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (uint32_t d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Block-wide exclusive scan of one u64 per thread (blockDim.x a multiple of
// 64, at most 8 waves per smem slot count); returns the block total.
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t x, uint64_t *excl, uint64_t *smem /*[nwaves]*/) {
  const uint32_t t = threadIdx.x, w = t / 64, l = t % 64;
  const uint32_t nw = blockDim.x / 64;
  uint64_t inc = wave_incl_scan_u64(x);
  if (l == 63) smem[w] = inc;
  __syncthreads();
  uint64_t wbase = 0, tot = 0;
  for (uint32_t i = 0; i < nw; ++i) {
    uint64_t s = smem[i];
    if (i < w) wbase += s;
    tot += s;
  }
  __syncthreads();
  *excl = wbase + inc - x;
  return tot;
}

// Largest r in [0, 64) with starts[r] <= item (starts ascending, starts[0] == 0).
__device__ __forceinline__ uint32_t seg_of(const uint32_t *starts, uint32_t item) {
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t step = 32; step >= 1; step >>= 1)
    if (starts[lo + step] <= item && lo + step < 64) lo += step;
  return lo;
}

// Binary search for v in sorted a[0..n) (device or mapped host memory); true
// if present.
template <class T>
__device__ __forceinline__ bool sorted_contains(const T *a, uint64_t n, T v) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    T x = ldg(a + mid);
    if (x < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && ldg(a + lo) == v;
}

// First index of sorted a[0..n) whose value is >= v (lower) or > v (upper).
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
  uint32_t lo = 0;
  while (n) {
    const uint32_t h = n >> 1;
    if (ldg(a + lo + h) < v) {
      lo += h + 1;
      n -= h + 1;
    } else {
      n = h;
    }
  }
  return lo;
}
__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
  uint32_t lo = 0;
  while (n) {
    const uint32_t h = n >> 1;
    if (ldg(a + lo + h) <= v) {
      lo += h + 1;
      n -= h + 1;
    } else {
      n = h;
    }
  }
  return lo;
}

// Two membership tests at once: is v0 in sorted a0[0..n0) and v1 in
// a1[0..n1)?  k0 / k1 enter as "still wanted" (a false one is not searched)
// and leave as the result.  The two lower-bound searches advance in lockstep
// so their loads are in flight together, and a hit is seen when it is probed
// (a lower bound i with a[i] == v is always probed: the range's right end
// only moves onto i by probing it), so no load follows the search.
// GLOBAL: the lists are in device or mapped host memory (else in LDS).
template <bool GLOBAL = true>
__device__ __forceinline__ void contains_pair(const uint32_t *a0, uint32_t n0, uint32_t v0, bool &k0,
                                              const uint32_t *a1, uint32_t n1, uint32_t v1, bool &k1) {
  uint32_t lo0 = 0, lo1 = 0;
  bool f0 = false, f1 = false;
  if (!k0) n0 = 0;
  if (!k1) n1 = 0;
  while (n0 | n1) {
    const uint32_t h0 = n0 >> 1, h1 = n1 >> 1;
    uint32_t x0 = 0, x1 = 0;
    if (n0) x0 = GLOBAL ? ldg(a0 + lo0 + h0) : a0[lo0 + h0];
    if (n1) x1 = GLOBAL ? ldg(a1 + lo1 + h1) : a1[lo1 + h1];
    if (n0) {
      if (x0 < v0) {
        lo0 += h0 + 1;
        n0 -= h0 + 1;
      } else {
        f0 |= x0 == v0;
        n0 = h0;
      }
    }
    if (n1) {
      if (x1 < v1) {
        lo1 += h1 + 1;
        n1 -= h1 + 1;
      } else {
        f1 |= x1 == v1;
        n1 = h1;
      }
    }
  }
  k0 = k0 && f0;
  k1 = k1 && f1;
}

// Four membership tests at once (contains_pair for four values): v[g] in
// sorted a[g][0..n[g]); k[g] enters as "still wanted" and leaves as the
// result.  The four searches advance in lockstep, so up to four loads of one
// lane are in flight at every stage.
__device__ __forceinline__ void contains_quad(const uint32_t *const a[4], const uint32_t n_in[4],
                                              const uint32_t v[4], bool k[4]) {
  uint32_t lo[4], n[4];
  bool f[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    lo[g] = 0;
    n[g] = k[g] ? n_in[g] : 0;
    f[g] = false;
  }
  while (n[0] | n[1] | n[2] | n[3]) {
    uint32_t x[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) x[g] = n[g] ? ldg(a[g] + lo[g] + (n[g] >> 1)) : 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      if (!n[g]) continue;
      const uint32_t h = n[g] >> 1;
      if (x[g] < v[g]) {
        lo[g] += h + 1;
        n[g] -= h + 1;
      } else {
        f[g] |= x[g] == v[g];
        n[g] = h;
      }
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) k[g] = k[g] && f[g];
}

// Map a hardware workgroup id to a logical one so that consecutive logical
// blocks run on the same XCD (workgroups are dispatched round-robin over the
// 8 XCDs); neighbouring row blocks then share one L2.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t hw, uint32_t nblocks) {
  constexpr uint32_t X = 8;
  const uint32_t per = nblocks / X;
  if (per == 0) return hw;
  const uint32_t full = per * X;
  if (hw >= full) return hw;
  return (hw % X) * per + hw / X;
}

}  // namespace gemd
