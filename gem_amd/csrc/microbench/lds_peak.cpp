// LDS read bandwidth of one MI355X, measured: every wave of a full grid
// reads 16 B per lane per step from its block's 64 KiB LDS array
// (ds_read_b128, consecutive lanes on consecutive 16-byte words: no bank
// conflicts) for a fixed number of steps; bytes read / kernel time (HIP
// events) is the LDS peak the batched path's LDS-staged kernels are held
// against.  Also times a device-to-device copy of 1 GiB (HBM read + write).
//
//   lds_peak [steps]      prints one JSON line
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                      \
      std::exit(1);                                                                     \
    }                                                                                   \
  } while (0)

constexpr int THREADS = 256;
constexpr int WORDS = 4096;  // uint4 words: 64 KiB per block

__global__ void __launch_bounds__(THREADS) k_lds_read(uint32_t steps, uint32_t *sink) {
  __shared__ uint4 buf[WORDS];
  for (int i = threadIdx.x; i < WORDS; i += THREADS) buf[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
  __syncthreads();
  uint4 acc = make_uint4(0, 0, 0, 0);
  uint32_t idx = threadIdx.x;
  for (uint32_t s = 0; s < steps; ++s) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint4 v = buf[(idx + u * THREADS) & (WORDS - 1)];
      acc.x ^= v.x;
      acc.y ^= v.y;
      acc.z ^= v.z;
      acc.w ^= v.w;
    }
    idx = (idx + 8 * THREADS) & (WORDS - 1);  // the next 8 words of this lane
  }
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u) sink[blockIdx.x] = acc.x;
}

int main(int argc, char **argv) {
  const uint32_t steps = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20000;
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = cus * 8;  // 64 KiB per block: two blocks per CU at a time; the rest queue
  uint32_t *sink = nullptr;
  CHECK(hipMalloc(&sink, sizeof(uint32_t) * blocks));
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  k_lds_read<<<blocks, THREADS>>>(10, sink);  // warm-up
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(a));
  k_lds_read<<<blocks, THREADS>>>(steps, sink);
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, a, b));
  const double bytes = (double)blocks * THREADS * 16.0 * 8.0 * steps;
  // HBM: 1 GiB device-to-device copy
  const size_t n = 1ull << 30;
  void *x = nullptr, *y = nullptr;
  CHECK(hipMalloc(&x, n));
  CHECK(hipMalloc(&y, n));
  CHECK(hipMemset(x, 1, n));
  CHECK(hipMemcpy(y, x, n, hipMemcpyDeviceToDevice));
  CHECK(hipEventRecord(a));
  for (int i = 0; i < 5; ++i) CHECK(hipMemcpy(y, x, n, hipMemcpyDeviceToDevice));
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float cms = 0;
  CHECK(hipEventElapsedTime(&cms, a, b));
  std::printf(
      "{\"cus\": %d, \"lds_blocks\": %d, \"lds_bytes\": %.0f, \"lds_ms\": %.3f, \"lds_TBps\": %.2f, "
      "\"copy_bytes_moved\": %.0f, \"copy_ms\": %.3f, \"copy_TBps\": %.2f}\n",
      cus, blocks, bytes, ms, bytes / (ms * 1e-3) / 1e12, 2.0 * 5 * n, cms, 2.0 * 5 * n / (cms * 1e-3) / 1e12);
  CHECK(hipFree(x));
  CHECK(hipFree(y));
  CHECK(hipFree(sink));
  return 0;
}
