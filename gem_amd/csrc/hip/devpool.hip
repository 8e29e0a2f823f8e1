// Caching device allocator behind DevBuf (declared in ctx.h): the bookkeeping
// of host/pool_core.h over hipMalloc / hipFree, shared by every context of
// the process and serialised by one mutex.
//
// The batched join allocates and frees a level buffer per running query per
// iteration, hundreds of thousands per batch, from 256 B to several GiB, and
// sized by a bound, so hardly two requests are alike.  Driver allocations
// (hipMalloc / hipMallocAsync) cost tens of microseconds each and far more for
// large sizes, so memory is taken from the driver in 8 GiB segments and small
// size classes and reused.  Parked small blocks, and empty segments beyond
// half of the device memory, go back to the driver at the end of every
// batched run (pool_trim); everything empty goes back when the driver refuses
// an allocation, and when a context is destroyed.
#include <mutex>

#include "../host/pool_core.h"
#include "ctx.h"

namespace gemd {

namespace {

void *hip_backing_alloc(size_t n, void *) {
  void *p = nullptr;
  if (hipMalloc(&p, n) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
void hip_backing_release(void *p, void *) { (void)hipFree(p); }
void hip_backing_sync(void *) { (void)hipDeviceSynchronize(); }

std::mutex g_mu;
gemh::PoolCore &core() {
  static gemh::PoolCore *c =
      new gemh::PoolCore(gemh::PoolBacking{hip_backing_alloc, hip_backing_release, hip_backing_sync, nullptr});
  return *c;  // never destroyed: buffers may be freed during process exit
}

}  // namespace

void *pool_alloc(size_t n, hipStream_t s, size_t *cls) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    return core().alloc(n, s, cls);
  } catch (const gemh::PoolCore::OutOfMemory &e) {
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const gemh::PoolStats st = core().stats();
    throw Error{GEM_E_NOMEM, "device allocation of " + std::to_string(e.bytes) + " bytes refused (pool: " +
                                 std::to_string(st.in_use) + " bytes in use, " + std::to_string(st.held) +
                                 " held; driver: " + std::to_string(free_b) + " free)"};
  }
}

void pool_free(void *p, size_t cls, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  core().free(p, cls, s);
}

size_t pool_reusable_bytes() {
  std::lock_guard<std::mutex> lk(g_mu);
  const gemh::PoolStats st = core().stats();
  return st.empty_segments + st.cached_small;
}

uint64_t pool_trim(uint64_t retain) {
  std::lock_guard<std::mutex> lk(g_mu);
  return core().trim(retain);
}

gemh::PoolStats pool_stats(bool reset_peak) {
  std::lock_guard<std::mutex> lk(g_mu);
  const gemh::PoolStats st = core().stats();
  if (reset_peak) core().reset_peak();
  return st;
}

void pool_release(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  core().release_stream(s);
}

}  // namespace gemd
