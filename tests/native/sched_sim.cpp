// CPU harness for the batched join's host scheduler (gem_amd/csrc/host/vsched.h)
// and the caching allocator's bookkeeping (gem_amd/csrc/host/pool_core.h).
//
// A simulated backend stands in for the GPU: a level's rows exist only as a
// count, each 64-row unit's output bound and survivor count are drawn from a
// hash of (buffer id, unit), with rare "hub" units whose bounds are ~10^4 times
// the rest, and buffers come from a PoolCore over fake device addresses (no
// memory is touched).  The program runs the scheduler over many queries for
// several passes on one pool and prints one JSON object per pass.
//
//   sched_sim run <queries> <budget_bytes> <passes> <seed> [max_active] [swap_bytes]
//                 [eager] [every]
// eager/every: queries from `eager` on are prepared elsewhere, in order, one
// per `every` extend steps (0 = all ready at the start), as a batch whose later
// queries are prepared on a second stream while the join runs.
//   sched_sim plan      (fixed plan_chunks cases, one JSON line)
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pool_core.h"
#include "vsched.h"

namespace {

uint64_t mix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

struct FakeDevice {
  uint64_t next = 1ull << 40;
  uint64_t capacity = 1ull << 42, held = 0;
};
void *fake_alloc(size_t n, void *u) {
  FakeDevice *d = static_cast<FakeDevice *>(u);
  if (d->held + n > d->capacity) return nullptr;
  void *p = reinterpret_cast<void *>(d->next);
  d->next += (n + 4095) / 4096 * 4096 + 4096;
  d->held += n;
  return p;
}
void fake_release(void *, void *) {}
void fake_sync(void *) {}

const void *const STREAM = reinterpret_cast<const void *>(0x1234);

struct SimBuf {
  gemh::PoolCore *pool = nullptr;
  void *p = nullptr;
  size_t cls = 0;
  uint64_t bytes = 0, id = 0;
  uint64_t nrows = 0;  // rows the buffer holds (its level's row count)
  ~SimBuf() {
    if (p) pool->free(p, cls, STREAM);
  }
};

struct SimBackend {
  using Buf = SimBuf;
  gemh::PoolCore *pool;
  uint64_t seed;
  uint64_t next_id = 1;
  uint64_t length_items = 0, extend_calls = 0;
  // lazy preparation: query q >= eager is ready once every * (q - eager + 1)
  // extend steps have run, or once the schedule waited for it (the preparer
  // works in order, so a wait for q readies every query up to q)
  const std::vector<gemh::SQuery<SimBuf> *> *qs = nullptr;
  uint32_t eager = ~0u;
  uint64_t every = 0;
  uint32_t waited_up_to = 0;  // queries below it are ready
  std::vector<char> started;
  uint64_t ready_waits = 0, not_ready = 0;
  bool is_ready(uint32_t q) const {
    return q < eager || q < waited_up_to || extend_calls >= every * (uint64_t)(q - eager + 1);
  }
  void check_ready(uint32_t q, const char *what) const {
    if (!is_ready(q)) {
      std::fprintf(stderr, "query %u %s before it was ready\n", q, what);
      std::exit(6);
    }
  }

  static uint64_t buf_bytes(const SimBuf &b) { return b.bytes; }

  std::shared_ptr<SimBuf> alloc(uint64_t bytes) {
    auto b = std::make_shared<SimBuf>();
    b->pool = pool;
    b->p = pool->alloc(bytes, STREAM, &b->cls);
    b->bytes = bytes;
    b->id = next_id++;
    return b;
  }
  std::shared_ptr<SimBuf> shrink(const std::shared_ptr<SimBuf> &from, uint64_t bytes) {
    auto b = alloc(bytes);
    b->id = from->id;  // same rows
    b->nrows = from->nrows;
    return b;
  }
  // output bound of unit u of buffer `id` holding `rows` rows at `level`
  uint64_t bound(uint64_t id, uint32_t level, uint64_t u, uint64_t nrows) const {
    const uint64_t r = std::min<uint64_t>(gemh::SCHED_UNIT, nrows - u * gemh::SCHED_UNIT);
    const uint64_t h = mix(seed ^ (id * 0x100000001b3ull) ^ (u << 8) ^ level);
    if (h % 1000 < 3) return r * (20000 + h % 30000);  // hub unit: long lists
    return r * (h % 9);                                 // ordinary unit: 0-8 per row
  }
  uint64_t survivors(uint64_t id, uint32_t level, uint64_t u, uint64_t nrows, bool final_level) const {
    const uint64_t b = bound(id, level, u, nrows);
    const uint64_t h = mix(seed ^ 0x5bd1e995ull ^ (id << 20) ^ u ^ ((uint64_t)level << 50));
    if (final_level) return b * (h % 101) / 100;
    if (b > 64 * 1000) return b * (h % 8) / 100;   // hub units: a few percent survive
    return b * (h % 31) / 100;                      // up to 30% survive: deep levels shrink
  }
  void length_pass(const std::vector<gemh::SItem<SimBuf>> &items, std::vector<std::vector<uint64_t>> *prefix) {
    prefix->resize(items.size());
    for (size_t t = 0; t < items.size(); ++t) {
      const auto &x = items[t];
      check_ready(x.q, "was measured");
      if (!x.rows) {
        std::fprintf(stderr, "a level in host memory was measured\n");
        std::exit(5);
      }
      const uint64_t nrows = x.row_end - x.row_begin;
      const uint64_t nu = (nrows + gemh::SCHED_UNIT - 1) / gemh::SCHED_UNIT;
      std::vector<uint64_t> &P = (*prefix)[t];
      P.assign(nu + 1, 0);
      for (uint64_t u = 0; u < nu; ++u) P[u + 1] = P[u] + bound(x.rows->id, x.level, u, nrows);
      ++length_items;
    }
  }
  uint64_t level_rows_of(const gemh::SItem<SimBuf> &x) const { return x.rows->nrows; }
  void extend(const std::vector<gemh::SItem<SimBuf>> &wr, const std::vector<gemh::SItem<SimBuf>> &fin,
              std::vector<uint64_t> *got_wr, std::vector<uint64_t> *got_fin) {
    ++extend_calls;
    got_wr->assign(wr.size(), 0);
    got_fin->assign(fin.size(), 0);
    for (size_t t = 0; t < wr.size(); ++t) {
      const auto &x = wr[t];
      check_ready(x.q, "was extended");
      if (!x.rows) {
        std::fprintf(stderr, "a level in host memory was extended\n");
        std::exit(5);
      }
      const uint64_t nrows = level_rows_of(x);
      for (uint64_t u = x.row_begin / gemh::SCHED_UNIT; u * gemh::SCHED_UNIT < x.row_end; ++u)
        (*got_wr)[t] += survivors(x.rows->id, x.level, u, nrows, false);
      if ((*got_wr)[t] > x.out_rows) {
        std::fprintf(stderr, "survivors above the bound\n");
        std::exit(3);
      }
      x.out->nrows = (*got_wr)[t];
    }
    for (size_t t = 0; t < fin.size(); ++t) {
      const auto &x = fin[t];
      check_ready(x.q, "was counted");
      const uint64_t nrows = level_rows_of(x);
      for (uint64_t u = x.row_begin / gemh::SCHED_UNIT; u * gemh::SCHED_UNIT < x.row_end; ++u)
        (*got_fin)[t] += survivors(x.rows->id, x.level, u, nrows, true);
    }
  }
  void on_iteration(uint64_t, size_t, uint64_t, uint64_t) {}
  bool restart(uint32_t) { return false; }
  bool prolong(uint32_t) { return false; }
  // a query the schedule is told is ready starts at once
  bool ready(uint32_t q) {
    if (!is_ready(q)) {
      ++not_ready;
      return false;
    }
    started[q] = 1;
    return true;
  }
  // waiting is for when nothing runs: no started query has levels pending
  void wait_ready(uint32_t q) {
    for (size_t i = 0; i < qs->size(); ++i)
      if (started[i] && !(*qs)[i]->done && !(*qs)[i]->stack.empty()) {
        std::fprintf(stderr, "waited for query %u while query %zu runs\n", q, i);
        std::exit(7);
      }
    ++ready_waits;
    waited_up_to = std::max(waited_up_to, q + 1);
    started[q] = 1;
  }
  void admit(uint32_t) {}
  // swapping: a host copy keeps the rows' identity (buffer id, rows)
  struct HostRows {
    uint64_t id, nrows, bytes;
  };
  uint64_t host_live = 0, host_peak = 0;
  std::shared_ptr<void> swap_out(const SimBuf &b, uint64_t bytes) {
    host_live += bytes;
    host_peak = std::max(host_peak, host_live);
    SimBackend *self = this;
    return std::shared_ptr<void>(new HostRows{b.id, b.nrows, bytes}, [self](void *p) {
      self->host_live -= static_cast<HostRows *>(p)->bytes;
      delete static_cast<HostRows *>(p);
    });
  }
  std::shared_ptr<SimBuf> swap_in(const std::shared_ptr<void> &h, uint64_t bytes) {
    const HostRows &r = *static_cast<const HostRows *>(h.get());
    if (r.bytes != bytes) {
      std::fprintf(stderr, "swapped level came back with another size\n");
      std::exit(4);
    }
    auto b = alloc(bytes);
    b->id = r.id;
    b->nrows = r.nrows;
    return b;
  }
};

int run(int nq, uint64_t budget, int passes, uint64_t seed, uint32_t max_active, uint64_t swap_bytes, uint32_t eager,
        uint64_t every) {
  FakeDevice dev;
  gemh::PoolCore pool(gemh::PoolBacking{fake_alloc, fake_release, fake_sync, &dev});
  for (int pass = 0; pass < passes; ++pass) {
    SimBackend be{&pool, seed + (uint64_t)pass * 7919};
    std::vector<gemh::SQuery<SimBuf>> qs(nq);
    std::vector<gemh::SQuery<SimBuf> *> ptrs;
    uint32_t max_nv = 0;
    for (int i = 0; i < nq; ++i) {
      const uint64_t h = mix(seed * 31 + (uint64_t)i);
      gemh::SQuery<SimBuf> &Q = qs[i];
      Q.nv = 4 + (uint32_t)(h % 29);
      max_nv = std::max(max_nv, Q.nv);
      Q.nb.assign(Q.nv, 1);
      for (uint32_t l = 1; l < Q.nv; ++l) Q.nb[l] = 1 + (uint32_t)(mix(h + l) % 3);
      Q.level_rows.assign(Q.nv, 0);
      const uint64_t n0 = 1000 + (h >> 8) % 200000;
      Q.level_rows[0] = n0;
      gemh::SLevel<SimBuf> L;
      L.level = 1;
      L.rows = be.alloc(4 * n0);
      L.rows->nrows = n0;
      L.nrows = n0;
      L.k = 1;
      Q.stack.push_back(std::move(L));
      ptrs.push_back(&Q);
    }
    if (every) {
      be.qs = &ptrs;
      be.eager = eager;
      be.every = every;
    }
    be.started.assign(nq, 0);
    gemh::SchedConfig cfg;
    cfg.cap = 1000000000ull;
    cfg.work_limit = 2000000000ull;
    cfg.budget = budget;
    cfg.max_chunk = 1ull << 30;
    cfg.max_active = max_active;
    cfg.swap_bytes = swap_bytes;
    gemh::SchedStats st;
    const gemh::PoolStats before = pool.stats();
    pool.reset_peak();
    gemh::run_schedule(be, ptrs, cfg, &st);
    uint64_t saturated = 0, limited = 0, finished = 0, pending = 0;
    for (auto &Q : qs) {
      saturated += Q.saturated;
      limited += Q.work_limited;
      finished += Q.done && !Q.saturated && !Q.work_limited;
      pending += !Q.stack.empty();
    }
    qs.clear();  // every level buffer goes back to the pool
    const gemh::PoolStats after = pool.stats();
    const uint64_t trimmed = pool.trim();
    const gemh::PoolStats trimmed_st = pool.stats();
    const uint64_t force_chunk = budget / 32;
    std::printf(
        "{\"pass\": %d, \"queries\": %d, \"max_nv\": %u, \"budget\": %" PRIu64 ", \"force_chunk\": %" PRIu64
        ", \"iterations\": %" PRIu64 ", \"forced\": %" PRIu64 ", \"waits\": %" PRIu64 ", \"peak_charged\": %" PRIu64
        ", \"max_overshoot\": %" PRIu64 ", \"max_unit_bytes\": %" PRIu64 ", \"pool_in_use_before\": %" PRIu64
        ", \"pool_peak\": %" PRIu64 ", \"pool_in_use_after\": %" PRIu64 ", \"pool_held_after\": %" PRIu64
        ", \"trimmed\": %" PRIu64 ", \"held_after_trim\": %" PRIu64 ", \"saturated\": %" PRIu64
        ", \"work_limited\": %" PRIu64 ", \"finished\": %" PRIu64 ", \"pending\": %" PRIu64
        ", \"length_items\": %" PRIu64 ", \"extend_calls\": %" PRIu64 ", \"swaps_out\": %" PRIu64
        ", \"swaps_in\": %" PRIu64 ", \"swap_bytes_out\": %" PRIu64 ", \"peak_host\": %" PRIu64
        ", \"backend_host_peak\": %" PRIu64 ", \"host_left\": %" PRIu64 ", \"ready_waits\": %" PRIu64
        ", \"not_ready\": %" PRIu64 ", \"started\": %" PRIu64 "}\n",
        pass, nq, max_nv, budget, force_chunk, st.iterations, st.forced, st.waits, st.peak_charged,
        st.max_overshoot, st.max_unit_bytes, before.in_use, after.peak_in_use, after.in_use, after.held, trimmed,
        trimmed_st.held, saturated, limited, finished, pending, be.length_items, be.extend_calls, st.swaps_out,
        st.swaps_in, st.swap_bytes_out, st.peak_host, be.host_peak, be.host_live, be.ready_waits, be.not_ready,
        (uint64_t)std::count(be.started.begin(), be.started.end(), 1));
    std::fflush(stdout);
  }
  return 0;
}

// Fixed plan_chunks cases: print the end units and the forced index.
int plan_cases() {
  auto show = [](const char *name, const std::vector<gemh::ChunkCand> &c, uint64_t left, uint64_t maxc,
                 uint64_t force, bool other) {
    std::vector<uint64_t> end;
    const int f = gemh::plan_chunks(c, left, maxc, force, other, &end);
    std::printf("\"%s\": {\"forced\": %d, \"end\": [", name, f);
    for (size_t i = 0; i < end.size(); ++i) std::printf("%s%" PRIu64, i ? ", " : "", end[i]);
    std::printf("]}");
  };
  // prefixes in rows; row_bytes 8; fields: prefix, cursor, nunits, row bytes,
  // depth, live, quota, levels left
  const std::vector<uint64_t> a = {0, 10, 20, 30, 40, 50};    // 10 rows (80 B) per unit
  const std::vector<uint64_t> b = {0, 100, 100, 100, 1100};   // a hub unit at the end
  const std::vector<uint64_t> z = {0, 0, 0, 5};               // zero-bound units, then 5
  std::printf("{");
  // quotas of 400 B: a (two levels to write) may use 200 B -> 2 units; b's
  // first unit (800 B) exceeds its whole free quota -> it waits
  show("share", {{&a, 0, 5, 8, 1, 0, 400, 2}, {&b, 0, 4, 8, 3, 0, 400, 1}}, 800, 1 << 20, 64, false);
  std::printf(", ");
  // nothing fits, nothing else runs: the deepest (b, depth 3) is forced one unit
  // (force chunk 64 B < its 800 B unit)
  show("forced", {{&a, 0, 5, 8, 1, 0, 400, 2}, {&b, 0, 4, 8, 3, 0, 400, 1}}, 0, 1 << 20, 64, false);
  std::printf(", ");
  // nothing fits but count-only work runs: everyone waits
  show("wait", {{&a, 0, 5, 8, 1, 0, 400, 2}, {&b, 0, 4, 8, 3, 0, 400, 1}}, 0, 1 << 20, 64, true);
  std::printf(", ");
  // zero-bound units cost nothing: taken even with no budget
  show("zero", {{&z, 0, 3, 8, 2, 0, 0, 1}}, 0, 1 << 20, 64, true);
  std::printf(", ");
  // forced with a large force chunk: b takes units 1-3 (100 rows = 800 B), not the hub
  show("force_chunk", {{&a, 0, 5, 8, 1, 0, 400, 2}, {&b, 0, 4, 8, 3, 0, 400, 1}}, 0, 1 << 20, 1000, false);
  std::printf(", ");
  // max_chunk caps a single query's share
  show("max_chunk", {{&a, 0, 5, 8, 1, 0, 1 << 20, 1}}, 1 << 20, 160, 64, false);
  std::printf(", ");
  // a share (100 / 4 levels = 25 B) below one unit (80 B) that fits the whole
  // free quota: one unit
  show("one_unit", {{&a, 0, 5, 8, 1, 0, 100, 4}}, 1 << 20, 1 << 20, 64, false);
  std::printf(", ");
  // pending rows count against the quota: 400 - 300 = 100 B -> one unit
  show("quota_live", {{&a, 0, 5, 8, 1, 300, 400, 1}}, 1 << 20, 1 << 20, 64, false);
  std::printf("}\n");
  return 0;
}

// Requests on both sides of the small-class / segment boundary (1 MiB), in
// two rounds with frees in between: in-use returns to zero and a trim returns
// everything to the backing allocator.
int pool_cases() {
  FakeDevice dev;
  gemh::PoolCore pool(gemh::PoolBacking{fake_alloc, fake_release, fake_sync, &dev});
  const size_t sizes[] = {1,      255,    256,    257,     4096,     900000,  917504,  950000,
                          983040, 1000000, 1048575, 1048576, 1048577, 1114112, 5000000, 9ull << 30};
  std::printf("{\"rounds\": [");
  for (int round = 0; round < 2; ++round) {
    std::vector<std::pair<void *, size_t>> live;
    uint64_t expect = 0;
    bool classes_ok = true;
    for (size_t n : sizes) {
      size_t cls = 0;
      void *p = pool.alloc(n, STREAM, &cls);
      classes_ok = classes_ok && cls >= n && (cls < (size_t(1) << 20)) == (n < 917505);
      live.push_back({p, cls});
      expect += cls;
    }
    const uint64_t in_use_full = pool.stats().in_use;
    for (auto &x : live) pool.free(x.first, x.second, STREAM);
    const gemh::PoolStats after = pool.stats();
    std::printf("%s{\"in_use_full\": %" PRIu64 ", \"expect\": %" PRIu64 ", \"in_use_after\": %" PRIu64
                ", \"classes_ok\": %s}",
                round ? ", " : "", in_use_full, expect, after.in_use, classes_ok ? "true" : "false");
  }
  // a trim that retains 10 GiB keeps one 8 GiB segment and some small
  // blocks, and holds at most that much; a second one keeps it all (nothing
  // to give back down to 10 GiB); a full trim returns the rest
  const uint64_t held_full = pool.stats().held;
  const uint64_t rel_retain = pool.trim(10ull << 30);
  const uint64_t held_retain = pool.stats().held;
  const uint64_t rel_again = pool.trim(10ull << 30);
  pool.trim();
  std::printf("], \"held_full\": %" PRIu64 ", \"released_retain\": %" PRIu64 ", \"held_retain\": %" PRIu64
              ", \"released_again\": %" PRIu64 ", \"held_after_trim\": %" PRIu64 "}\n",
              held_full, rel_retain, held_retain, rel_again, pool.stats().held);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && !std::strcmp(argv[1], "plan")) return plan_cases();
  if (argc >= 2 && !std::strcmp(argv[1], "pool")) return pool_cases();
  if (argc >= 6 && !std::strcmp(argv[1], "run"))
    return run(std::atoi(argv[2]), std::strtoull(argv[3], nullptr, 10), std::atoi(argv[4]),
               std::strtoull(argv[5], nullptr, 10), argc >= 7 ? (uint32_t)std::atoi(argv[6]) : 0u,
               argc >= 8 ? std::strtoull(argv[7], nullptr, 10) : 0ull,
               argc >= 9 ? (uint32_t)std::atoi(argv[8]) : 0u, argc >= 10 ? std::strtoull(argv[9], nullptr, 10) : 0ull);
  std::fprintf(stderr, "usage: sched_sim run <queries> <budget> <passes> <seed> [max_active] | sched_sim plan\n");
  return 2;
}
