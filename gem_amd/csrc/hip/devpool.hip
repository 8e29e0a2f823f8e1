// Caching device allocator behind DevBuf (declared in ctx.h).
//
// The batched join allocates and frees a level buffer per running query per
// iteration, hundreds of thousands per batch, from 256 B to several GiB, and
// sized by a bound, so hardly two requests are alike.  Driver allocations
// (hipMalloc / hipMallocAsync) cost tens of microseconds each and far more for
// large sizes, so memory is taken from the driver in big pieces and reused:
//   * requests under 1 MiB: size classes (four per power of two), freed blocks
//     parked on a free list per (stream, class);
//   * larger requests: 8 GiB segments (or one segment of the request's size
//     when larger) carved best-fit, in 64 KiB units, with freed ranges merged
//     with their free neighbours.
// Memory freed by one stream is only reused by that stream: all device work of
// a context runs on its one stream, so a range the host frees after
// enqueueing its last use can serve later work there without a wait.  When
// the driver refuses more memory, every cached block and every empty segment
// is returned to it (after a device synchronisation) and the request retried.
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "ctx.h"

namespace gemd {

namespace {

constexpr size_t BIG = size_t(1) << 20;      // requests from this size go to segments
constexpr size_t UNIT = size_t(64) << 10;    // segment carving granularity
constexpr size_t SEG = size_t(8) << 30;      // segment size

struct Segment {
  char *base = nullptr;
  size_t size = 0, used = 0;
  hipStream_t s = nullptr;
  std::map<size_t, size_t> by_off;                 // free ranges: offset -> length
  std::multimap<size_t, size_t> by_len;            // the same ranges: length -> offset

  void add_free(size_t off, size_t len) {
    by_off.emplace(off, len);
    by_len.emplace(len, off);
  }
  void erase_free(std::map<size_t, size_t>::iterator it) {
    auto r = by_len.equal_range(it->second);
    for (auto j = r.first; j != r.second; ++j)
      if (j->second == it->first) {
        by_len.erase(j);
        break;
      }
    by_off.erase(it);
  }
};

std::mutex g_mu;
std::map<std::pair<hipStream_t, size_t>, std::vector<void *>> g_small;  // (stream, class) -> free blocks
std::map<char *, Segment> g_segs;                                       // by base address

size_t small_class(size_t n) {
  if (n <= 256) return 256;
  const int e = 63 - __builtin_clzll((unsigned long long)(n - 1));
  const size_t base = size_t(1) << e, step = base >> 2;
  return base + (n - base + step - 1) / step * step;
}

// Return every cached small block and every empty segment to the driver;
// caller holds g_mu.
void drop_cached_locked() {
  (void)hipDeviceSynchronize();
  for (auto &kv : g_small)
    for (void *p : kv.second) (void)hipFree(p);
  g_small.clear();
  for (auto it = g_segs.begin(); it != g_segs.end();) {
    if (it->second.used == 0) {
      (void)hipFree(it->second.base);
      it = g_segs.erase(it);
    } else {
      ++it;
    }
  }
}

void *driver_alloc_locked(size_t n) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    drop_cached_locked();
    e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error{GEM_E_NOMEM, "hipMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e)};
    }
  }
  return p;
}

// Best fit over the stream's segments; nullptr when none has room.
char *carve_locked(hipStream_t s, size_t n) {
  Segment *best = nullptr;
  std::multimap<size_t, size_t>::iterator bit;
  for (auto &kv : g_segs) {
    Segment &g = kv.second;
    if (g.s != s) continue;
    auto it = g.by_len.lower_bound(n);
    if (it != g.by_len.end() && (!best || it->first < bit->first)) {
      best = &g;
      bit = it;
    }
  }
  if (!best) return nullptr;
  const size_t off = bit->second, len = bit->first;
  best->erase_free(best->by_off.find(off));
  if (len > n) best->add_free(off + n, len - n);
  best->used += n;
  return best->base + off;
}

}  // namespace

void *pool_alloc(size_t n, hipStream_t s, size_t *cls) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (n < BIG) {
    const size_t c = small_class(n);
    *cls = c;
    auto it = g_small.find({s, c});
    if (it != g_small.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      return p;
    }
    return driver_alloc_locked(c);
  }
  const size_t c = (n + UNIT - 1) / UNIT * UNIT;
  *cls = c;
  if (char *p = carve_locked(s, c)) return p;
  size_t sz = c > SEG ? c : SEG;
  char *base = nullptr;
  try {
    base = static_cast<char *>(driver_alloc_locked(sz));
  } catch (const Error &) {
    if (sz == c) throw;
    sz = c;  // no room for a whole segment: one of the request's size
    base = static_cast<char *>(driver_alloc_locked(sz));
  }
  Segment &g = g_segs[base];
  g.base = base;
  g.size = sz;
  g.s = s;
  g.add_free(0, sz);
  return carve_locked(s, c);
}

void pool_free(void *ptr, size_t cls, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (cls < BIG) {
    g_small[{s, cls}].push_back(ptr);
    return;
  }
  char *p = static_cast<char *>(ptr);
  auto sit = g_segs.upper_bound(p);
  if (sit == g_segs.begin()) return;  // its segment was released with its stream
  --sit;
  Segment &g = sit->second;
  if (g.s != s || p >= g.base + g.size) return;
  size_t off = (size_t)(p - g.base), len = cls;
  g.used -= len;
  // merge with the free neighbours
  auto nx = g.by_off.lower_bound(off);
  if (nx != g.by_off.end() && nx->first == off + len) {
    len += nx->second;
    g.erase_free(nx);
  }
  auto pv = g.by_off.lower_bound(off);
  if (pv != g.by_off.begin()) {
    --pv;
    if (pv->first + pv->second == off) {
      off = pv->first;
      len += pv->second;
      g.erase_free(pv);
    }
  }
  g.add_free(off, len);
}

size_t pool_cached_bytes() {
  std::lock_guard<std::mutex> lk(g_mu);
  size_t n = 0;
  for (auto &kv : g_small) n += kv.first.second * kv.second.size();
  for (auto &kv : g_segs) n += kv.second.size - kv.second.used;
  return n;
}

void pool_release(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto it = g_small.begin(); it != g_small.end();) {
    if (it->first.first == s) {
      for (void *p : it->second) (void)hipFree(p);
      it = g_small.erase(it);
    } else {
      ++it;
    }
  }
  for (auto it = g_segs.begin(); it != g_segs.end();) {
    if (it->second.s == s) {
      (void)hipFree(it->second.base);
      it = g_segs.erase(it);
    } else {
      ++it;
    }
  }
}

}  // namespace gemd
