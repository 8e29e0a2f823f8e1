// Host-side graph types and helpers shared by the C ABI, the planner and the
// CLI.  Pure C++17; no HIP.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

struct gem_graph {
  uint32_t n = 0;
  std::vector<uint64_t> off;  // n + 1
  std::vector<uint32_t> nbr;  // off[n]
  std::vector<uint32_t> lab;  // n
};

namespace gemh {

// splitmix64 (Steele, Lea, Flood 2014): a counter-based mixer, so sample i of a
// generator can be recomputed independently of any other sample and of the
// thread that draws it.
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t x = s;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  }
  // uniform double in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, n) (multiply-shift; bias < n / 2^64)
  uint64_t below(uint64_t n) {
    return (uint64_t)(((unsigned __int128)next() * (unsigned __int128)n) >> 64);
  }
};

inline unsigned hw_threads() {
  unsigned t = std::thread::hardware_concurrency();
  return t ? t : 1;
}

// Run f(begin, end) over [0, n) split into contiguous chunks, one per thread.
template <class F>
void parallel_chunks(uint64_t n, F f, unsigned threads = 0) {
  if (threads == 0) threads = hw_threads();
  if (n < 4096 || threads == 1) {
    f((uint64_t)0, n);
    return;
  }
  std::vector<std::thread> ts;
  uint64_t chunk = (n + threads - 1) / threads;
  for (unsigned t = 0; t < threads; ++t) {
    uint64_t b = t * chunk, e = std::min<uint64_t>(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([=]() { f(b, e); });
  }
  for (auto &th : ts) th.join();
}

// Run f(begin, end) over [0, n) in chunks of `grain` handed out from a shared
// counter (for skewed per-item work, e.g. the hub lists of R-MAT graphs).
template <class F>
void parallel_dynamic(uint64_t n, uint64_t grain, F f, unsigned threads = 0) {
  if (threads == 0) threads = hw_threads();
  if (n <= grain || threads == 1) {
    f((uint64_t)0, n);
    return;
  }
  std::atomic<uint64_t> next{0};
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < threads; ++t)
    ts.emplace_back([&]() {
      for (;;) {
        const uint64_t b = next.fetch_add(grain);
        if (b >= n) break;
        f(b, std::min<uint64_t>(n, b + grain));
      }
    });
  for (auto &th : ts) th.join();
}

// Build a symmetric, sorted, duplicate-free CSR from a generator of directed
// pairs.  `emit(i, &s, &d)` returns false to drop sample i; each kept sample
// contributes s->d and d->s.  The generator is called twice per sample (count
// pass, then scatter pass), so it must be a pure function of i.
template <class Emit>
void build_symmetric_csr(uint32_t n, uint64_t num_samples, Emit emit, gem_graph &g) {
  std::vector<uint32_t> deg(n, 0);
  parallel_chunks(num_samples, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      uint32_t s, d;
      if (!emit(i, &s, &d) || s == d) continue;
      __atomic_fetch_add(&deg[s], 1u, __ATOMIC_RELAXED);
      __atomic_fetch_add(&deg[d], 1u, __ATOMIC_RELAXED);
    }
  });
  std::vector<uint64_t> off(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) off[v + 1] = off[v] + deg[v];
  std::vector<uint32_t> adj(off[n]);
  std::vector<uint32_t> cur(n, 0);
  parallel_chunks(num_samples, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      uint32_t s, d;
      if (!emit(i, &s, &d) || s == d) continue;
      uint32_t ps = __atomic_fetch_add(&cur[s], 1u, __ATOMIC_RELAXED);
      adj[off[s] + ps] = d;
      uint32_t pd = __atomic_fetch_add(&cur[d], 1u, __ATOMIC_RELAXED);
      adj[off[d] + pd] = s;
    }
  });
  cur.clear();
  cur.shrink_to_fit();
  // sort + dedup each list (deterministic whatever order the scatter produced)
  parallel_dynamic(n, 1024, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; ++v) {
      uint32_t *p = adj.data() + off[v];
      uint32_t len = deg[v];
      std::sort(p, p + len);
      uint32_t w = 0;
      for (uint32_t i = 0; i < len; ++i)
        if (w == 0 || p[i] != p[w - 1]) p[w++] = p[i];
      deg[v] = w;
    }
  });
  g.n = n;
  g.off.assign(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) g.off[v + 1] = g.off[v] + deg[v];
  g.nbr.resize(g.off[n]);
  parallel_chunks(n, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; ++v)
      std::copy(adj.begin() + off[v], adj.begin() + off[v] + deg[v], g.nbr.begin() + g.off[v]);
  });
}

// Check: sorted ascending, no self loops, no duplicates, symmetric.
bool validate_csr(const gem_graph &g, std::string *why);

}  // namespace gemh
