// Host scheduler of the batched vertex-at-a-time join: per query a depth-first
// stack of pending level ranges, the length passes that bound each level's
// output, chunk planning against a shared memory budget, and the bookkeeping
// of counts, caps and work limits.  Independent of HIP: the device work (and
// its buffers) is behind a backend, so the same code runs in the library and
// in CPU tests with a simulated backend.
//
// Memory rule.  Pending level rows of every query ("live") plus the output
// buffers allocated in an iteration are charged against `budget`.  Each
// running query has a quota: budget / max_active while queries wait to start,
// budget / (queries running) once none waits, but at most tail_growth times
// the first quota (quotas only grow then).  A
// query's next chunk may use (quota - its pending bytes) / (levels it still
// has to write) bytes, so the levels after it find room too and a deep query
// keeps descending in sizeable chunks instead of the budget filling up with
// shallow levels.  A query whose next 64-row unit does not fit that share but
// fits its whole free quota takes that one unit; otherwise it waits.  Only
// when no query can take anything and no count-only work is pending is the
// budget exceeded: the deepest pending level (whose memory is released
// soonest) takes a chunk of at most `force_chunk` bytes, or one unit if that
// is larger.  Such forced chunks pile up along one depth-first path at most,
// so the charged bytes never exceed
//   budget + (levels - 1) * max(force_chunk, largest unit bound).
//
// Restarts.  A query may carry a limit on the work of its current attempt
// (attempt_limit).  When an attempt goes past it, its pending levels are
// dropped and the backend may start the query again (another vertex order, a
// larger limit); the final count does not depend on the order, so whichever
// attempt ends gives the query's result.  Before that, the backend may prolong
// the attempt instead (raise its limit), when what it counted so far says it
// is near its end.
//
// Swapping (swap_bytes != 0).  When the device-resident pending rows exceed
// the budget after an extend step, the least recently used pending levels
// that are not on top of a stack move to host memory (up to swap_bytes in
// all), and a level comes back to the device when it is on top of its stack
// at the start of an iteration.  Depth-first order makes the least recently
// used levels the ones needed last: the shallow levels under a query's
// current path.  The rows are the same wherever they are, so counts do not
// depend on it.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

namespace gemh {

// Rows per planning unit (one wave task of the extend kernels).
constexpr uint32_t SCHED_UNIT = 64;

struct SchedConfig {
  uint64_t cap = 0;         // count cap (0 = none): a query stops once its count exceeds it
  uint64_t work_limit = 0;  // 0 = none
  uint64_t budget = 0;      // bytes for pending + newly allocated level rows
  uint64_t max_chunk = 4ull << 30;       // largest output buffer of one query in one iteration
  uint64_t force_chunk = 0;              // past the budget (see above); 0 = budget / 32
  uint64_t final_piece = 1ull << 26;     // least output bound of a capped final-level piece
  uint64_t shrink_min = 4ull << 20;      // outputs at least this large move when half unused
  // queries running at once (0 = all): the others wait in index order and
  // start as running ones end (continuous admission; a query's first level
  // is seeded when it starts)
  uint32_t max_active = 0;
  uint32_t tail_growth = 8;  // most a quota grows once no query waits (see run_schedule)
  // queries [0, primaries) are the batch; later ones (0 = none) race a
  // partner among them, start only once every primary has started, and do
  // not count as waiting for the quotas
  uint32_t primaries = 0;
  uint64_t swap_bytes = 0;  // host memory pending levels may move to (0 = none)
};

struct SchedStats {
  uint64_t iterations = 0;
  uint64_t length_passes = 0;
  uint64_t forced = 0;          // iterations that forced a unit past the budget
  uint64_t waits = 0;           // (query, iteration) pairs deferred for lack of budget
  uint64_t peak_charged = 0;    // max over iterations of live + this iteration's outputs
  uint64_t max_overshoot = 0;   // max over iterations of charged bytes beyond the budget
  uint64_t max_unit_bytes = 0;  // largest output bound of a forced unit
  uint64_t extend_items = 0;    // items handed to extend steps
  uint64_t restarts = 0;        // attempts ended at their limit and started again
  uint64_t admitted = 0;        // queries started
  uint64_t swaps_out = 0, swaps_in = 0;            // levels moved to host memory / back
  uint64_t swap_bytes_out = 0, swap_bytes_in = 0;  // their bytes
  uint64_t peak_host = 0;                          // most bytes held in host memory at once
};

template <class Buf>
struct SLevel {
  uint32_t level = 0;  // the level these rows extend to next (query vertex vorder[level])
  std::shared_ptr<Buf> rows;
  uint64_t nrows = 0;
  uint32_t k = 0;                 // columns per row
  std::vector<uint64_t> unit_u;   // prefix of per-unit output bounds (nunits + 1) once measured
  uint64_t cursor = 0;            // next unit
  bool measured = false;
  // swapping: the rows in host memory (rows null meanwhile), their bytes, and
  // the iteration that last pushed or read this level
  std::shared_ptr<void> host;
  uint64_t host_bytes = 0;
  uint64_t last_use = 0;
  uint64_t nunits() const { return (nrows + SCHED_UNIT - 1) / SCHED_UNIT; }
};

template <class Buf>
struct SQuery {
  uint32_t nv = 0;
  std::vector<uint32_t> nb;  // backward lists of each level (nb[i], i >= 1)
  std::vector<SLevel<Buf>> stack;
  std::vector<uint64_t> level_rows;  // rows after each level
  uint64_t count = 0, processed = 0;
  // rows read plus (upper bound of) list entries walked: what the work limit
  // bounds, since one row with hub lists can cost 10^5 entries
  uint64_t work = 0;
  bool saturated = false, done = false, work_limited = false;
  // current attempt: work at its start and the work it may add (0 = no
  // limit); attempts = restarts so far
  uint64_t attempt_start = 0, attempt_limit = 0;
  uint32_t attempts = 0;
  // measured final-level pieces of the current attempt: output bounds and
  // embeddings counted (their ratio sizes the next capped piece)
  uint64_t fin_bound = 0, fin_got = 0;
  // Racing: another query of the schedule computing the same count (another
  // vertex order).  Whichever ends first with a count cancels the other.
  int64_t partner = -1;
  bool cancelled = false;
};

// One unit of a launch: rows [row_begin, row_end) of a pending level of
// query `q`, extended to `level`; `out` (write items) receives at most
// out_rows rows of k + 1 columns.
template <class Buf>
struct SItem {
  uint32_t q = 0, level = 0, k = 0;
  const Buf *rows = nullptr;
  uint64_t row_begin = 0, row_end = 0;
  Buf *out = nullptr;
  uint64_t out_rows = 0;
};

// Chunk planning of one iteration.  cand[i]: a query whose top level writes
// output; prefix = its unit bound prefix, cursor = next unit, row_bytes =
// bytes of one output row, depth = its level, live = bytes of its pending
// levels, quota = what they may hold, levels_left = levels it still has to
// write, this one included.  `left` = budget not yet charged; other_progress
// = count-only work runs this iteration anyway.  Returns the end unit of each
// query's chunk (== cursor: it waits) and the index of the query forced past
// the budget (-1 if none).
struct ChunkCand {
  const std::vector<uint64_t> *prefix;
  uint64_t cursor, nunits, row_bytes;
  uint32_t depth;
  uint64_t live, quota;
  uint32_t levels_left;
};
inline int plan_chunks(const std::vector<ChunkCand> &cand, uint64_t left, uint64_t max_chunk, uint64_t force_chunk,
                       bool other_progress, std::vector<uint64_t> *end) {
  const size_t n = cand.size();
  end->resize(n);
  for (size_t i = 0; i < n; ++i) (*end)[i] = cand[i].cursor;
  if (!n) return -1;
  auto bytes = [&](size_t i, uint64_t e) {
    const std::vector<uint64_t> &P = *cand[i].prefix;
    return (P[e] - P[cand[i].cursor]) * cand[i].row_bytes;
  };
  // largest e in [cursor, nunits] with bytes(i, e) <= lim
  auto reach = [&](size_t i, uint64_t lim) {
    uint64_t lo = cand[i].cursor, hi = cand[i].nunits;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo + 1) / 2;
      if (bytes(i, mid) <= lim)
        lo = mid;
      else
        hi = mid - 1;
    }
    return lo;
  };
  // deepest levels first: their memory is released soonest
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cand[a].depth > cand[b].depth; });
  uint64_t room = left;
  for (size_t i : order) {
    const ChunkCand &c = cand[i];
    const uint64_t free_q = c.quota > c.live ? c.quota - c.live : 0;
    const uint64_t share = std::min<uint64_t>({max_chunk, free_q / std::max<uint32_t>(c.levels_left, 1), room});
    uint64_t e = reach(i, share);
    if (e == c.cursor && c.cursor < c.nunits) {
      const uint64_t one = bytes(i, c.cursor + 1);
      if (one <= std::min<uint64_t>({free_q, room, max_chunk})) e = c.cursor + 1;
    }
    (*end)[i] = e;
    room -= bytes(i, e);
  }
  bool any = other_progress;
  for (size_t i = 0; i < n && !any; ++i) any = (*end)[i] > cand[i].cursor;
  if (any) return -1;
  // nothing fits and nothing else moves: the deepest level goes past the
  // budget by one chunk (ties: the smaller next unit)
  size_t f = 0;
  for (size_t i = 1; i < n; ++i) {
    if (cand[i].depth > cand[f].depth ||
        (cand[i].depth == cand[f].depth && bytes(i, cand[i].cursor + 1) < bytes(f, cand[f].cursor + 1)))
      f = i;
  }
  (*end)[f] = std::max<uint64_t>(cand[f].cursor + 1, reach(f, std::min(max_chunk, force_chunk)));
  return (int)f;
}

// Backend interface (a template parameter, not a virtual base):
//   using Buf = ...;
//   static uint64_t buf_bytes(const Buf &);
//   std::shared_ptr<Buf> alloc(uint64_t bytes);
//   // per item, the prefix of per-unit output bounds over its rows
//   void length_pass(const std::vector<SItem<Buf>> &items, std::vector<std::vector<uint64_t>> *prefix);
//   // rows written per write item, embeddings counted per final item
//   void extend(const std::vector<SItem<Buf>> &wr, const std::vector<SItem<Buf>> &fin,
//               std::vector<uint64_t> *got_wr, std::vector<uint64_t> *got_fin);
//   // a buffer of `bytes` holding the first `bytes` of `from`
//   std::shared_ptr<Buf> shrink(const std::shared_ptr<Buf> &from, uint64_t bytes);
//   void on_iteration(uint64_t iter, size_t active, uint64_t live, uint64_t budget);
//   // query q's attempt went past its limit and its stack is empty: seed it
//   // again for a new attempt (count, level_rows, attempt fields) and return
//   // true, or return false to end it unfinished
//   bool restart(uint32_t q);
//   // query q's attempt went past its limit, its levels still pending: raise
//   // its attempt_limit and return true to let it go on, or return false
//   bool prolong(uint32_t q);
//   // is query q ready to start (its candidates and tables prepared)?  If not,
//   // the schedule starts it later, or waits for it when nothing else runs
//   bool ready(uint32_t q);
//   void wait_ready(uint32_t q);
//   // swapping only: copy `bytes` of a device buffer to new host memory, and
//   // back to a new device buffer (copies ordered on the backend's stream)
//   std::shared_ptr<void> swap_out(const Buf &rows, uint64_t bytes);
//   std::shared_ptr<Buf> swap_in(const std::shared_ptr<void> &host, uint64_t bytes);
//   // query q starts: seed its first level (or mark it done); only called
//   // for queries whose stack is empty and that are not done
//   void admit(uint32_t q);
//
// The queries' nb / nv must be filled; a query whose stack is already seeded
// (level 1 = the identity rows of the first vertex's candidates) when the
// schedule starts is not passed to admit().
template <class BE>
void run_schedule(BE &be, std::vector<SQuery<typename BE::Buf> *> &qs, const SchedConfig &cfg, SchedStats *st) {
  using Buf = typename BE::Buf;
  using Item = SItem<Buf>;
  const uint64_t budget = cfg.budget;
  auto live_bytes = [&]() {
    uint64_t b = 0;
    for (SQuery<Buf> *S : qs)
      for (const SLevel<Buf> &L : S->stack) b += L.rows ? BE::buf_bytes(*L.rows) : 0;
    return b;
  };
  std::vector<uint8_t> started(qs.size(), 0);
  size_t next = 0;
  uint64_t host_used = 0;  // swapped levels' bytes
  const size_t nprim = cfg.primaries ? std::min<size_t>(cfg.primaries, qs.size()) : qs.size();
  // a query that ends with a count (or saturated) cancels its partner
  auto cancel_partner = [&](SQuery<Buf> &S) {
    if (S.partner < 0) return;
    SQuery<Buf> &P = *qs[(size_t)S.partner];
    if (P.done) return;
    P.done = true;
    P.cancelled = true;
    P.stack.clear();
  };
  for (;;) {
    // ---- admission (index order)
    {
      size_t running = 0;
      for (uint32_t i = 0; i < qs.size(); ++i)
        if (started[i] && !qs[i]->done && !qs[i]->stack.empty()) ++running;
      while (next < qs.size() && (!cfg.max_active || running < cfg.max_active)) {
        const uint32_t qi = (uint32_t)next;
        // a query prepared elsewhere (another stream) starts once it is
        // ready: later, while others run, or now, after waiting for it
        if (!qs[qi]->done && !be.ready(qi)) {
          if (running) break;
          be.wait_ready(qi);
        }
        ++next;
        started[qi] = 1;
        if (qs[qi]->done) continue;
        if (qs[qi]->stack.empty()) be.admit(qi);
        if (!qs[qi]->done && !qs[qi]->stack.empty()) {
          ++running;
          ++st->admitted;
        }
      }
    }
    std::vector<uint32_t> act;
    for (uint32_t i = 0; i < qs.size(); ++i)
      if (started[i] && !qs[i]->done && !qs[i]->stack.empty()) act.push_back(i);

    if (act.empty()) break;
    ++st->iterations;
    // swapping: every running query's top level is on the device
    if (cfg.swap_bytes)
      for (uint32_t qi : act) {
        SLevel<Buf> &top = qs[qi]->stack.back();
        top.last_use = st->iterations;
        if (top.host) {
          top.rows = be.swap_in(top.host, top.host_bytes);
          host_used -= top.host_bytes;
          ++st->swaps_in;
          st->swap_bytes_in += top.host_bytes;
          top.host.reset();
          top.host_bytes = 0;
        }
      }
    const uint64_t live = live_bytes();
    be.on_iteration(st->iterations, act.size(), live, budget);

    // ---- length pass: output bounds of every unmeasured top level (final
    // levels only when they walk lists under a cap: their units are then
    // counted a cap-sized piece at a time)
    {
      std::vector<Item> items;
      std::vector<uint32_t> owner;
      for (uint32_t qi : act) {
        SQuery<Buf> &S = *qs[qi];
        SLevel<Buf> &top = S.stack.back();
        const bool final_level = top.level + 1 == S.nv;
        if (top.measured || (final_level && (S.nb[top.level] == 1 || !cfg.cap))) continue;
        Item it;
        it.q = qi;
        it.level = top.level;
        it.k = top.k;
        it.rows = top.rows.get();
        it.row_begin = 0;
        it.row_end = top.nrows;
        items.push_back(it);
        owner.push_back(qi);
      }
      if (!items.empty()) {
        ++st->length_passes;
        std::vector<std::vector<uint64_t>> pref;
        be.length_pass(items, &pref);
        for (size_t t = 0; t < owner.size(); ++t) {
          SLevel<Buf> &top = qs[owner[t]]->stack.back();
          top.unit_u = std::move(pref[t]);
          top.measured = true;
        }
      }
    }

    // ---- count-only pieces of final levels (no output memory)
    std::vector<Item> wr, fin;
    std::vector<std::shared_ptr<Buf>> keep_alive, outs;
    std::vector<uint32_t> wr_owner, fin_owner;
    std::vector<uint64_t> fin_bounds;  // output bound of each measured final piece (0: unmeasured)
    std::vector<ChunkCand> cand;
    std::vector<uint32_t> cand_owner;
    // quotas: budget / max_active while queries wait to start; then the
    // budget over the queries still running, at most tail_growth times the
    // first quota (the last queries take larger chunks, so they do not crawl
    // through thousands of small iterations, but a capped query given the
    // whole budget would walk its first levels so broadly that it can do
    // many times the work it needs to reach the cap)
    uint64_t shares = act.size();
    if (cfg.max_active)
      shares = next < nprim ? cfg.max_active
                            : std::max<uint64_t>(shares, cfg.max_active / std::max<uint32_t>(cfg.tail_growth, 1));
    const uint64_t quota = budget / std::max<uint64_t>(shares, 1);
    for (uint32_t qi : act) {
      SQuery<Buf> &S = *qs[qi];
      SLevel<Buf> &top = S.stack.back();
      const uint64_t nu = top.nunits();
      if (top.level + 1 != S.nv) {
        uint64_t mine = 0;
        for (const SLevel<Buf> &L : S.stack) mine += L.rows ? BE::buf_bytes(*L.rows) : 0;
        cand.push_back(ChunkCand{&top.unit_u, top.cursor, nu, 4ull * (top.k + 1), top.level, mine, quota,
                                 S.nv - 1 - top.level});
        cand_owner.push_back(qi);
        continue;
      }
      uint64_t b1 = nu, piece_bound = 0;
      if (top.measured) {
        // a piece whose output bound should just exceed the cap: what is
        // still missing, over the share of bounds that earlier pieces turned
        // into embeddings (1 before any), plus 10% (at least final_piece
        // list entries of work)
        const uint64_t need = cfg.cap + 1 > S.count ? cfg.cap + 1 - S.count : 1;
        const double ratio = S.fin_bound >= (1ull << 20) && S.fin_got
                                 ? std::min(1.0, (double)S.fin_got / (double)S.fin_bound)
                                 : 1.0;
        const double want = 1.1 * (double)need / ratio;
        const uint64_t target = std::max<uint64_t>(want >= 1.8e19 ? ~0ull : (uint64_t)want, cfg.final_piece);
        const std::vector<uint64_t> &P = top.unit_u;
        uint64_t lo = top.cursor + 1, hi = nu;  // smallest b1 with P[b1] - P[cursor] >= target, else nu
        while (lo < hi) {
          const uint64_t mid = lo + (hi - lo) / 2;
          if (P[mid] - P[top.cursor] >= target)
            hi = mid;
          else
            lo = mid + 1;
        }
        b1 = lo;
        piece_bound = P[b1] - P[top.cursor];
        S.work += piece_bound;
      }
      Item it;
      it.q = qi;
      it.level = top.level;
      it.k = top.k;
      it.rows = top.rows.get();
      it.row_begin = top.cursor * SCHED_UNIT;
      it.row_end = std::min<uint64_t>(b1 * SCHED_UNIT, top.nrows);
      keep_alive.push_back(top.rows);
      S.processed += it.row_end - it.row_begin;
      S.work += it.row_end - it.row_begin;
      top.cursor = b1;
      fin.push_back(it);
      fin_owner.push_back(qi);
      fin_bounds.push_back(piece_bound);
      if (top.cursor >= nu) S.stack.pop_back();
    }

    // ---- output chunks under the budget
    const uint64_t left = budget > live ? budget - live : 0;
    std::vector<uint64_t> end;
    const int forced = plan_chunks(cand, left, cfg.max_chunk, cfg.force_chunk ? cfg.force_chunk : budget / 32,
                                   !fin.empty(), &end);
    uint64_t charged = live;
    for (size_t c = 0; c < cand.size(); ++c) {
      const uint32_t qi = cand_owner[c];
      SQuery<Buf> &S = *qs[qi];
      SLevel<Buf> &top = S.stack.back();
      const uint64_t b0 = top.cursor, b1 = end[c];
      if (b1 == b0) {
        ++st->waits;
        continue;
      }
      const uint64_t ub = top.unit_u[b1] - top.unit_u[b0];
      const uint64_t rb = 4ull * (top.k + 1);
      Item it;
      it.q = qi;
      it.level = top.level;
      it.k = top.k;
      it.rows = top.rows.get();
      it.row_begin = b0 * SCHED_UNIT;
      it.row_end = std::min<uint64_t>(b1 * SCHED_UNIT, top.nrows);
      S.processed += it.row_end - it.row_begin;
      S.work += it.row_end - it.row_begin + ub;
      top.cursor = b1;
      if (ub) {
        keep_alive.push_back(top.rows);
        std::shared_ptr<Buf> out = be.alloc(ub * rb);
        charged += ub * rb;
        it.out = out.get();
        it.out_rows = ub;
        outs.push_back(std::move(out));
        wr.push_back(it);
        wr_owner.push_back(qi);
        if ((int)c == forced) st->max_unit_bytes = std::max(st->max_unit_bytes, ub * rb);
      }
      if (top.cursor >= top.nunits()) S.stack.pop_back();
    }
    if (forced >= 0) ++st->forced;
    st->peak_charged = std::max(st->peak_charged, charged);
    if (charged > budget) st->max_overshoot = std::max(st->max_overshoot, charged - budget);

    // ---- extend step
    if (!wr.empty() || !fin.empty()) {
      st->extend_items += wr.size() + fin.size();
      std::vector<uint64_t> got_wr, got_fin;
      be.extend(wr, fin, &got_wr, &got_fin);
      for (size_t t = 0; t < wr.size(); ++t) {
        SQuery<Buf> &S = *qs[wr_owner[t]];
        const uint64_t got = got_wr[t];
        const Item &it = wr[t];
        S.level_rows[it.level] += got;
        std::shared_ptr<Buf> rows = std::move(outs[t]);
        if (!got) continue;  // the buffer goes back now
        // the output was allocated for the length pass's bound; when fewer
        // than half of those rows survived, they move to a buffer of their
        // size, and the bound-sized one goes back at once (pending rows are
        // what the budget is charged with)
        const uint64_t rb = 4ull * (it.k + 1);
        if (2 * got <= it.out_rows && it.out_rows * rb >= cfg.shrink_min) rows = be.shrink(rows, got * rb);
        SLevel<Buf> L;
        L.level = it.level + 1;
        L.rows = std::move(rows);
        L.nrows = got;
        L.k = it.k + 1;
        L.last_use = st->iterations;
        S.stack.push_back(std::move(L));
      }
      for (size_t t = 0; t < fin.size(); ++t) {
        SQuery<Buf> &S = *qs[fin_owner[t]];
        S.level_rows[S.nv - 1] += got_fin[t];
        S.count += got_fin[t];
        if (fin_bounds[t]) {
          S.fin_bound += fin_bounds[t];
          S.fin_got += got_fin[t];
        }
      }
    }
    for (uint32_t qi : act) {
      SQuery<Buf> &S = *qs[qi];
      if (S.done) continue;  // cancelled by its partner this iteration
      if (cfg.cap && S.count > cfg.cap) {
        S.saturated = true;
        S.done = true;
        S.stack.clear();
        cancel_partner(S);
      } else if (cfg.work_limit && S.work > cfg.work_limit) {
        S.work_limited = true;
        S.done = true;
        S.stack.clear();
      } else if (S.attempt_limit && S.work - S.attempt_start > S.attempt_limit && !be.prolong(qi)) {
        S.stack.clear();
        if (be.restart(qi)) {
          ++st->restarts;
        } else {
          S.work_limited = true;
          S.done = true;
        }
      } else if (S.stack.empty()) {
        S.done = true;
        cancel_partner(S);
      }
    }
    // swapping: device-resident pending rows above the budget go to host
    // memory, least recently used first, never a level on top of its stack
    if (cfg.swap_bytes) {
      uint64_t dev = live_bytes();
      if (dev > budget) {
        std::vector<SLevel<Buf> *> cands;
        for (SQuery<Buf> *S : qs)
          for (size_t i = 0; i + 1 < S->stack.size(); ++i)
            if (S->stack[i].rows) cands.push_back(&S->stack[i]);
        std::stable_sort(cands.begin(), cands.end(),
                         [](const SLevel<Buf> *a, const SLevel<Buf> *b) { return a->last_use < b->last_use; });
        for (SLevel<Buf> *L : cands) {
          if (dev <= budget) break;
          const uint64_t b = BE::buf_bytes(*L->rows);
          if (host_used + b > cfg.swap_bytes) continue;
          L->host = be.swap_out(*L->rows, b);
          L->host_bytes = b;
          L->rows.reset();
          host_used += b;
          dev -= b;
          ++st->swaps_out;
          st->swap_bytes_out += b;
        }
      }
      // a cleared stack (saturated, cancelled, restarted) has released its host copies
      uint64_t h = 0;
      for (SQuery<Buf> *S : qs)
        for (const SLevel<Buf> &L : S->stack) h += L.host ? L.host_bytes : 0;
      host_used = h;
      st->peak_host = std::max(st->peak_host, host_used);
    }
  }
}

}  // namespace gemh
