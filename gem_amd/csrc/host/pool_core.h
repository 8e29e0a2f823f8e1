// Bookkeeping of the library's caching device allocator, independent of HIP:
// the backing allocator (hipMalloc / hipFree on the device, anything on the
// host) is passed in, so the same code runs in the library and in CPU tests.
//
//   * requests under `big` bytes: size classes (four per power of two, 256 B
//     at least), freed blocks parked on a free list per (stream, class);
//   * larger requests: segments of `seg` bytes (or one of the request's size
//     when larger, or when a whole segment is refused) carved best-fit in
//     `unit`-byte units, freed ranges merged with their free neighbours.
// Memory freed by one stream is only reused by that stream (all device work
// of a context runs on its one stream, so a range freed after its last use
// was enqueued can serve later work there without a wait).  When the backing
// allocator refuses memory, every cached block and every empty segment is
// returned to it and the request retried once.
//
// Not thread-safe: the owner serialises calls.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace gemh {

struct PoolBacking {
  void *(*alloc)(size_t n, void *user);  // nullptr when refused
  void (*release)(void *p, void *user);
  void (*sync)(void *user);  // wait for all work before memory goes back
  void *user;
};

struct PoolStats {
  uint64_t in_use = 0;       // bytes of live allocations (rounded to their class)
  uint64_t peak_in_use = 0;  // highest in_use since the last reset_peak()
  uint64_t held = 0;         // bytes obtained from the backing allocator
  uint64_t cached_small = 0; // parked small blocks
  uint64_t empty_segments = 0;  // bytes of segments with nothing carved
  uint64_t segments = 0;
  uint64_t backing_allocs = 0;  // calls to the backing allocator
};

class PoolCore {
 public:
  struct OutOfMemory {
    size_t bytes;
  };

  PoolCore(PoolBacking b, size_t big = size_t(1) << 20, size_t unit = size_t(64) << 10,
           size_t seg = size_t(8) << 30)
      : b_(b), big_(big), unit_(unit), seg_(seg) {}
  ~PoolCore() { release_all(); }
  PoolCore(const PoolCore &) = delete;
  PoolCore &operator=(const PoolCore &) = delete;

  static size_t small_class(size_t n) {
    if (n <= 256) return 256;
    const int e = 63 - __builtin_clzll((unsigned long long)(n - 1));
    const size_t base = size_t(1) << e, step = base >> 2;
    return base + (n - base + step - 1) / step * step;
  }

  // Throws OutOfMemory when the backing allocator refuses even after the
  // cache was dropped.  *cls receives the block size (pass it to free()).
  void *alloc(size_t n, const void *stream, size_t *cls) {
    // the class decides the path (free() tells the two apart by it): a
    // request just under `big` whose class rounds up to `big` is carved
    if (n < big_ && small_class(n) < big_) {
      const size_t c = small_class(n);
      *cls = c;
      void *p = nullptr;
      auto it = small_.find({stream, c});
      if (it != small_.end() && !it->second.empty()) {
        p = it->second.back();
        it->second.pop_back();
        st_.cached_small -= c;
      } else {
        p = backing_alloc(c);
        if (!p) throw OutOfMemory{c};
        st_.held += c;
      }
      note_use(c);
      return p;
    }
    const size_t c = std::max(big_, (n + unit_ - 1) / unit_ * unit_);
    *cls = c;
    char *p = carve(stream, c);
    if (!p) {
      size_t sz = c > seg_ ? c : seg_;
      char *base = static_cast<char *>(backing_alloc(sz));
      if (!base && sz != c) {
        sz = c;  // no room for a whole segment: one of the request's size
        base = static_cast<char *>(backing_alloc(sz));
      }
      if (!base) throw OutOfMemory{c};
      st_.held += sz;
      Segment &g = segs_[base];
      g.base = base;
      g.size = sz;
      g.stream = stream;
      g.add_free(0, sz);
      p = carve(stream, c);
    }
    note_use(c);
    return p;
  }

  void free(void *ptr, size_t cls, const void *stream) {
    if (!ptr) return;
    if (cls < big_) {
      st_.in_use -= cls;
      small_[{stream, cls}].push_back(ptr);
      st_.cached_small += cls;
      return;
    }
    char *p = static_cast<char *>(ptr);
    auto sit = segs_.upper_bound(p);
    if (sit == segs_.begin()) return;  // its segment was released with its stream
    --sit;
    Segment &g = sit->second;
    if (g.stream != stream || p >= g.base + g.size) return;
    st_.in_use -= cls;
    size_t off = (size_t)(p - g.base), len = cls;
    g.used -= len;
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

  // Return parked small blocks and empty segments to the backing allocator
  // (after waiting for all work), as long as the memory held exceeds
  // `retain` bytes: segments first, then small blocks.  Returns the bytes
  // released.
  uint64_t trim(uint64_t retain = 0) {
    if (small_.empty() && !has_empty_segment()) return 0;
    if (st_.held <= retain) return 0;
    if (b_.sync) b_.sync(b_.user);
    uint64_t n = 0;
    for (auto it = segs_.begin(); it != segs_.end();) {
      if (it->second.used == 0 && st_.held - n > retain) {
        b_.release(it->second.base, b_.user);
        n += it->second.size;
        it = segs_.erase(it);
      } else {
        ++it;
      }
    }
    for (auto it = small_.begin(); it != small_.end() && st_.held - n > retain;) {
      std::vector<void *> &v = it->second;
      while (!v.empty() && st_.held - n > retain) {
        b_.release(v.back(), b_.user);
        v.pop_back();
        n += it->first.second;
        st_.cached_small -= it->first.second;
      }
      it = v.empty() ? small_.erase(it) : std::next(it);
    }
    st_.held -= n;
    return n;
  }

  // Release everything that belongs to `stream` (its work must be finished
  // and its buffers no longer used).
  void release_stream(const void *stream) {
    uint64_t n = 0;
    for (auto it = small_.begin(); it != small_.end();) {
      if (it->first.first == stream) {
        for (void *p : it->second) {
          b_.release(p, b_.user);
          n += it->first.second;
          st_.cached_small -= it->first.second;
        }
        it = small_.erase(it);
      } else {
        ++it;
      }
    }
    for (auto it = segs_.begin(); it != segs_.end();) {
      if (it->second.stream == stream) {
        st_.in_use -= it->second.used;
        b_.release(it->second.base, b_.user);
        n += it->second.size;
        it = segs_.erase(it);
      } else {
        ++it;
      }
    }
    st_.held -= n;
  }

  void release_all() {
    if (b_.sync) b_.sync(b_.user);
    for (auto &kv : small_)
      for (void *p : kv.second) b_.release(p, b_.user);
    small_.clear();
    for (auto &kv : segs_) b_.release(kv.second.base, b_.user);
    segs_.clear();
    st_ = PoolStats{};
  }

  PoolStats stats() const {
    PoolStats s = st_;
    s.empty_segments = 0;
    s.segments = segs_.size();
    for (auto &kv : segs_)
      if (kv.second.used == 0) s.empty_segments += kv.second.size;
    return s;
  }
  void reset_peak() { st_.peak_in_use = st_.in_use; }

 private:
  struct Segment {
    char *base = nullptr;
    size_t size = 0, used = 0;
    const void *stream = nullptr;
    std::map<size_t, size_t> by_off;       // free ranges: offset -> length
    std::multimap<size_t, size_t> by_len;  // the same ranges: length -> offset
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

  void note_use(size_t c) {
    st_.in_use += c;
    if (st_.in_use > st_.peak_in_use) st_.peak_in_use = st_.in_use;
  }
  bool has_empty_segment() const {
    for (auto &kv : segs_)
      if (kv.second.used == 0) return true;
    return false;
  }
  void *backing_alloc(size_t n) {
    ++st_.backing_allocs;
    void *p = b_.alloc(n, b_.user);
    if (p) return p;
    trim();
    ++st_.backing_allocs;
    return b_.alloc(n, b_.user);
  }
  char *carve(const void *stream, size_t n) {
    Segment *best = nullptr;
    std::multimap<size_t, size_t>::iterator bit;
    for (auto &kv : segs_) {
      Segment &g = kv.second;
      if (g.stream != stream) continue;
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

  PoolBacking b_;
  size_t big_, unit_, seg_;
  std::map<std::pair<const void *, size_t>, std::vector<void *>> small_;
  std::map<char *, Segment> segs_;
  PoolStats st_;
};

}  // namespace gemh
