// Host orchestration of the batched path: per group of queries, filter and
// edge tables for each query, a vertex plan per query (given, or derived from
// GEM's edge order), then one VBatchRunner over the whole group.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "ctx.h"
#include "tcache.h"
#include "vjoin.h"

namespace gemd {

static uint64_t pool_retain_bytes() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  return total_b / 2;
}

// Vertex order of first appearance along GEM's edge order.
static std::vector<uint32_t> gem_vertex_order(const QueryDev &Qd) {
  const gemh::QueryInfo &Q = Qd.Q;
  std::vector<uint64_t> sz(Q.ne);
  for (uint32_t e = 0; e < Q.ne; ++e) {
    const TableDev &a = Qd.tab[2 * e], &b = Qd.tab[2 * e + 1];
    sz[e] = a.present ? a.size : b.size;
  }
  std::vector<uint32_t> order(Q.ne), vorder;
  int st = gemh::gem_order(Q, sz.data(), order.data());
  if (st) throw Error{st, "GEM ordering failed"};
  st = gemh::vertex_order_from_edges(Q, order.data(), &vorder);
  if (st) throw Error{st, "vertex order from edge order failed"};
  return vorder;
}

static gemh::QueryStats query_stats(const gem_ctx *c, const QueryDev &Qd) {
  const gemh::QueryInfo &Q = Qd.Q;
  gemh::QueryStats s;
  s.ncan.assign(Qd.ncan.begin(), Qd.ncan.end());
  s.tsize.assign(2 * Q.ne, 0);
  for (uint32_t d = 0; d < 2 * Q.ne; ++d) s.tsize[d] = Qd.tab[d].present ? Qd.tab[d].size : 0;
  s.label_freq.assign(Q.nv, 0);
  for (uint32_t u = 0; u < Q.nv; ++u) {
    auto it = graph(c).label_id.find(Q.lab[u]);
    if (it != graph(c).label_id.end()) s.label_freq[u] = graph(c).label_freq[it->second];
  }
  s.data_v = graph(c).n;
  s.data_e = graph(c).slots / 2;
  return s;
}

static void plan_query(gem_ctx *c, VQueryState &S, const QueryDev &Qd, const uint32_t *vorder_in) {
  const gemh::QueryInfo &Q = Qd.Q;
  std::vector<uint32_t> vorder;
  if (vorder_in)
    vorder.assign(vorder_in, vorder_in + Q.nv);
  else
    vorder = gem_vertex_order(Qd);
  S.qd = &Qd;
  replan(c, S, vorder);
}

// Seed of a query's candidate orders: a hash of its labels and edges, so the
// candidates do not depend on where the query sits in a batch.
static uint64_t query_seed(const gemh::QueryInfo &Q) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ ((uint64_t)Q.nv << 32) ^ Q.ne;
  auto mix = [&](uint64_t x) {
    h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
  };
  for (uint32_t u = 0; u < Q.nv; ++u) mix(Q.lab[u]);
  for (const auto &e : Q.edges) mix(((uint64_t)e.first << 32) | e.second);
  return h;
}

static void fill_result(const VQueryState &V, uint64_t cap, gem_join_result *r, uint64_t *level_rows) {
  const gemh::SQuery<DevBuf> &S = V.s;
  r->count = S.work_limited ? 0 : (S.saturated ? cap : S.count);
  r->saturated = S.saturated ? 1 : 0;
  r->num_steps = V.qd->Q.nv - 1;
  r->num_embeddings_written = 0;
  r->rows_processed = S.processed;
  r->work_limited = S.work_limited ? 1 : 0;
  r->restarts = S.attempts;
  r->work = S.work;
  if (level_rows)
    for (uint32_t i = 0; i < GEM_MAX_QV; ++i) level_rows[i] = i < S.level_rows.size() ? S.level_rows[i] : 0;
}

void run_match_batch(gem_ctx *c, const gem_query_view *qs, uint32_t nq, const uint32_t *const *vorders,
                     const gem_batch_opts *opts, gem_join_result *results, uint64_t *level_rows) {
  if (!c->g.loaded) throw Error{GEM_E_STATE, "no data graph loaded"};
  gem_batch_opts o{};
  if (opts) o = *opts;
  const uint32_t group = o.group ? o.group : std::max<uint32_t>(nq, 1);
  const bool streamed = c->stream_cache_bytes != 0;
  hipStream_t s = c->stream;
  GEM_HIP_CHECK(hipEventRecord(c->ev[6], s));
  for (uint32_t g0 = 0; g0 < nq; g0 += group) {
    const uint32_t gn = std::min(group, nq - g0);
    const uint64_t budget = o.level_budget;  // shared by the running queries (0 = from free memory)
    std::vector<QueryDev> qd(gn);
    std::vector<VQueryState> st(gn);
    std::vector<VQueryState *> ptrs;
    for (uint32_t i = 0; i < gn; ++i) {
      run_filter(c, qd[i], &qs[g0 + i], 0);
      run_tables(c, qd[i]);
      plan_query(c, st[i], qd[i], vorders ? vorders[g0 + i] : nullptr);
      if (streamed) spill_tables(c, qd[i], c->stream_slot_bytes);
      ptrs.push_back(&st[i]);
    }
    VBatchRunner runner(c, o.cap, o.work_limit, budget, streamed);
    runner.run(ptrs);
    for (uint32_t i = 0; i < gn; ++i)
      fill_result(st[i], o.cap, &results[g0 + i], level_rows ? level_rows + (uint64_t)(g0 + i) * GEM_MAX_QV : nullptr);
  }
  GEM_HIP_CHECK(hipEventRecord(c->ev[7], s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  GEM_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
  c->times.join_ms = ms;
  prof_flush(c);
  // the run's level buffers and tables are gone: empty segments beyond half
  // of the device memory go back to the driver (the rest stays for the next
  // run, whose first allocations would otherwise wait for the driver)
  pool_trim(pool_retain_bytes());
}

void run_order_candidates(gem_ctx *c, uint32_t max_orders, uint64_t seed, uint32_t *orders, int32_t *feats,
                          uint32_t *n) {
  const QueryDev &Qd = c->q;
  if (!Qd.tabled) throw Error{GEM_E_STATE, "edge tables have not been built"};
  const gemh::QueryStats st = query_stats(c, Qd);
  std::vector<std::vector<uint32_t>> cand;
  // seed 0: the query's own hash, as the batched path seeds it
  gemh::candidate_orders(Qd.Q, st, gem_vertex_order(Qd), max_orders, seed ? seed : query_seed(Qd.Q), &cand);
  for (size_t i = 0; i < cand.size(); ++i) {
    if (orders) std::copy(cand[i].begin(), cand[i].end(), orders + i * Qd.Q.nv);
    if (feats) gemh::order_features(Qd.Q, st, cand[i], feats + i * gemh::FEAT_DIM);
  }
  *n = (uint32_t)cand.size();
}

// One query of a batched learned run: the device state whose candidates and
// tables it reads and, for a relaxed query that shares its parent's state,
// its own edges (it is planned on them) and the parent edge of each.
struct BatchItem {
  const QueryDev *qd = nullptr;
  const gemh::QueryInfo *Qp = nullptr;               // null: qd->Q
  const std::vector<uint32_t> *edge_map = nullptr;   // item edge -> qd edge
};

static const gemh::QueryInfo &item_query(const BatchItem &it) { return it.Qp ? *it.Qp : it.qd->Q; }

static const TableDev &item_table(const BatchItem &it, uint32_t d) {
  return it.qd->tab[it.edge_map ? 2 * (*it.edge_map)[d >> 1] + (d & 1) : d];
}

static std::vector<uint32_t> item_gem_vertex_order(const BatchItem &it) {
  const gemh::QueryInfo &Q = item_query(it);
  std::vector<uint64_t> sz(Q.ne);
  for (uint32_t e = 0; e < Q.ne; ++e) {
    const TableDev &a = item_table(it, 2 * e), &b = item_table(it, 2 * e + 1);
    sz[e] = a.present ? a.size : b.size;
  }
  std::vector<uint32_t> order(Q.ne), vorder;
  int st = gemh::gem_order(Q, sz.data(), order.data());
  if (st) throw Error{st, "GEM ordering failed"};
  st = gemh::vertex_order_from_edges(Q, order.data(), &vorder);
  if (st) throw Error{st, "vertex order from edge order failed"};
  return vorder;
}

static gemh::QueryStats item_stats(const gem_ctx *c, const BatchItem &it) {
  const gemh::QueryInfo &Q = item_query(it);
  gemh::QueryStats s = query_stats(c, *it.qd);
  s.tsize.assign(2 * Q.ne, 0);
  for (uint32_t d = 0; d < 2 * Q.ne; ++d) {
    const TableDev &T = item_table(it, d);
    s.tsize[d] = T.present ? T.size : 0;
  }
  return s;
}

// Upper bounds of a query's launch needs (see VQueryState::pending): the most
// backward lists of a level (the largest vertex degree) and equal-label
// columns (the most vertices of one label, less one).
static void launch_bounds_of(const gem_query_view &v, uint32_t *nb, uint32_t *nd) {
  std::vector<uint32_t> deg(v.num_v, 0);
  for (uint32_t e = 0; e < v.num_e; ++e) {
    deg[v.edges[2 * e]]++;
    deg[v.edges[2 * e + 1]]++;
  }
  std::map<uint32_t, uint32_t> per_label;
  uint32_t most = 0;
  for (uint32_t u = 0; u < v.num_v; ++u) most = std::max(most, ++per_label[v.labels[u]]);
  *nb = std::max<uint32_t>(1, deg.empty() ? 1 : *std::max_element(deg.begin(), deg.end()));
  *nd = most ? most - 1 : 0;
}

// One batched learned run over `items`: each query's candidate orders and
// features (once its filter and tables exist), their ranking and the plan of
// the first ranked order.  Queries prepared before the run are completed
// before it starts; the others are prepared on a second stream while the run
// goes on (mark_prepared) and completed when they start (LazyStates).
class ItemRun : public LazyStates {
 public:
  ItemRun(gem_ctx *c, std::vector<BatchItem> &items, gem_mlp *mlp, uint32_t max_orders, const gem_batch_opts &o)
      : c_(c), items_(items), mlp_(mlp), max_orders_(max_orders), o_(o), n_(items.size()), st_(items.size()),
        cand_(items.size()), feats_(items.size()), prepared_(items.size()) {
    for (auto &p : prepared_) p.store(false);
  }

  // Host part of query i's preparation, after its tables exist; cx is the
  // context that prepared it (any thread).
  void candidates(const gem_ctx *cx, size_t i) {
    const BatchItem &it = items_[i];
    const gemh::QueryInfo &Q = item_query(it);
    const gemh::QueryStats qsx = item_stats(cx, it);
    // without a model, candidates beyond GEM's order matter only to restarts
    const uint32_t norders = (mlp_ || o_.first_attempt) ? max_orders_ : 1;
    gemh::candidate_orders(Q, qsx, item_gem_vertex_order(it), norders, query_seed(Q), &cand_[i]);
    feats_[i].clear();
    if (mlp_) {
      feats_[i].resize(cand_[i].size() * gemh::FEAT_DIM);
      for (size_t j = 0; j < cand_[i].size(); ++j)
        gemh::order_features(Q, qsx, cand_[i][j], feats_[i].data() + j * gemh::FEAT_DIM);
    }
  }

  // Ranking of query i's candidates and the plan of the first (the run's
  // stream, the calling thread).  Without a model, as generated (GEM's order
  // first); with one, by predicted cost, except that GEM's order (candidate
  // 0) stays first unless the cheapest prediction beats it by more than the
  // model's margin.  The first ranked order runs; restarts take the next ones.
  void finish(size_t i) {
    std::vector<uint32_t> R(cand_[i].size());
    for (uint32_t j = 0; j < R.size(); ++j) R[j] = j;
    if (mlp_ && !R.empty()) {
      std::vector<float> p(R.size());
      const int rc = gem_mlp_predict(mlp_, feats_[i].data(), (uint32_t)R.size(), p.data());
      if (rc) throw Error{rc, "order model: " + c_->err};
      std::stable_sort(R.begin(), R.end(), [&](uint32_t a, uint32_t b) { return p[a] < p[b]; });
      if (!(p[R[0]] < p[0] - mlp_margin(mlp_))) {
        R.erase(std::find(R.begin(), R.end(), 0u));
        R.insert(R.begin(), 0u);
      }
    }
    VQueryState &V = st_[i];
    V.first_attempt = o_.first_attempt;
    V.growth = o_.restart_growth ? o_.restart_growth : 4;
    V.width = o_.restart_width ? o_.restart_width : 1;
    V.max_attempts = o_.max_attempts;
    V.ranked.clear();
    V.ranked_idx.clear();
    for (uint32_t j : R) {
      V.ranked.push_back(cand_[i][j]);
      V.ranked_idx.push_back(j);
    }
    // size the launch buffers for every order a restart may run
    if (attempt_limit(V, 0))
      for (size_t j = 1; j < V.ranked.size(); ++j) {
        gemh::VertexPlan P;
        if (gemh::make_vertex_plan(item_query(items_[i]), V.ranked[j], &P))
          throw Error{GEM_E_STATE, "invalid candidate vertex order"};
        for (size_t l = 1; l < P.levels.size(); ++l) {
          V.mnb_all = std::max(V.mnb_all, P.levels[l].nb);
          V.mnd_all = std::max(V.mnd_all, P.levels[l].ndup);
        }
      }
    replan(c_, V, V.ranked[0]);
    V.cur = 0;
  }

  void mark_prepared(size_t i) {
    std::lock_guard<std::mutex> lk(mu_);
    prepared_[i].store(true);
    cv_.notify_all();
  }
  void fail(std::exception_ptr e) {
    std::lock_guard<std::mutex> lk(mu_);
    error_ = e;
    cv_.notify_all();
  }
  bool ready(uint32_t q) override { return q < n_ && prepared_[q].load(); }
  void complete(uint32_t q) override {
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&]() { return prepared_[q].load() || error_; });
      if (!prepared_[q].load() && error_) std::rethrow_exception(error_);
    }
    if (st_[q].pending) finish(q);
  }

  // Runs the batch: queries [0, eager) are prepared; the rest are pending
  // until mark_prepared.  Results (and level_rows, chosen) indexed like items.
  void run(size_t eager, bool streamed, const std::vector<const gem_query_view *> &views,
           gem_join_result *results, uint64_t *level_rows, uint32_t *chosen) {
    std::vector<VQueryState *> ptrs;
    for (size_t i = 0; i < n_; ++i) {
      VQueryState &V = st_[i];
      V.qd = items_[i].qd;
      V.Qp = items_[i].Qp;
      V.edge_map = items_[i].edge_map;
      if (i < eager) {
        finish(i);
      } else {
        uint32_t nb = 1, nd = 0;
        launch_bounds_of(*views[i], &nb, &nd);
        V.pending = true;
        V.pending_nv = views[i]->num_v;
        V.max_nb = {nb};
        V.max_ndup = {nd};
      }
      ptrs.push_back(&V);
    }
    // Racing: each query with restarts gets a twin on the same candidates and
    // tables; the twins start, in the same order, once every query of the
    // batch has started, when a running query ends and leaves room (the tail,
    // where the few queries left would leave most of the device idle), and
    // whichever of a pair ends first with a count cancels the other.  Counts
    // do not depend on the order, so either gives the query's result.  A twin
    // takes its orders from its query when it starts.
    std::vector<VQueryState> twins;
    std::vector<int64_t> twin_of(n_, -1);
    if (c_->racing && !streamed && o_.first_attempt && max_orders_ > 1) {
      twins.resize(n_);
      for (size_t i = 0; i < n_; ++i) {
        VQueryState &P = st_[i];
        if (!P.pending && P.ranked.size() < 2) continue;
        VQueryState &T = twins[i];
        T.qd = P.qd;
        T.Qp = P.Qp;
        T.edge_map = P.edge_map;
        T.twin = true;
        T.pending = true;
        T.pending_nv = views[i]->num_v;
        if (P.pending) {
          T.max_nb = P.max_nb;
          T.max_ndup = P.max_ndup;
        } else {
          T.max_nb = {P.mnb_all};
          T.max_ndup = {P.mnd_all};
        }
        twin_of[i] = (int64_t)ptrs.size();
        T.s.partner = (int64_t)i;
        P.s.partner = twin_of[i];
        ptrs.push_back(&T);
      }
    }
    VBatchRunner runner(c_, o_.cap, o_.work_limit, o_.level_budget, streamed);
    if (!twins.empty()) runner.set_primaries((uint32_t)n_);
    runner.set_lazy(this);
    runner.run(ptrs);
    for (size_t i = 0; i < n_; ++i) {
      const VQueryState *W = &st_[i];
      uint64_t twin_work = 0;
      if (twin_of[i] >= 0) {
        const VQueryState &T = twins[i];
        twin_work = T.s.work;
        // the twin's result when it ended with a count and the query did not
        if (!T.pending && T.s.done && !T.s.cancelled && !T.s.work_limited &&
            (st_[i].s.cancelled || st_[i].s.work_limited))
          W = &T;
      }
      fill_result(*W, o_.cap, &results[i], level_rows ? level_rows + i * GEM_MAX_QV : nullptr);
      results[i].work = st_[i].s.work + twin_work;
      if (chosen) chosen[i] = W->ranked.empty() ? 0 : W->ranked_idx[W->cur];
    }
  }

 private:
  gem_ctx *c_;
  std::vector<BatchItem> &items_;
  gem_mlp *mlp_;
  uint32_t max_orders_;
  gem_batch_opts o_;
  size_t n_;
  std::vector<VQueryState> st_;
  std::vector<std::vector<std::vector<uint32_t>>> cand_;
  std::vector<std::vector<int32_t>> feats_;
  std::vector<std::atomic<bool>> prepared_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::exception_ptr error_;
};

// The context that prepares queries for c on a stream of its own (stream,
// events, staging), reading c's data graph; made at its first use.  The
// queries it prepares live in its stream's part of the device pool.
static gem_ctx *prep_ctx(gem_ctx *c) {
  if (!c->prep) {
    gem_ctx *p = nullptr;
    const int st = gem_hip_create(c->device, &p);
    if (st) throw Error{st, "cannot create the preparing context"};
    p->gsrc = &c->g;
    c->prep = p;
  }
  return c->prep;
}

// Prepares queries [begin, end) of a batch (filter, tables, candidate orders)
// on the preparing context in a thread of its own while the batch runs; stops
// after the query at hand and joins when destroyed (before the queries it
// writes are destroyed).
class Preparer {
 public:
  Preparer(gem_ctx *c, gem_ctx *shadow, const std::vector<const gem_query_view *> &views, std::vector<QueryDev> &qd,
           ItemRun &run, size_t begin, size_t end)
      : c_(c), shadow_(shadow) {
    th_ = std::thread([this, &views, &qd, &run, begin, end]() {
      try {
        GEM_HIP_CHECK(hipSetDevice(c_->device));
        using clk = std::chrono::steady_clock;
        for (size_t i = begin; i < end && !stop_.load(); ++i) {
          const auto ta = clk::now();
          run_filter(shadow_, qd[i], views[i], 0);
          const auto tb = clk::now();
          run_tables(shadow_, qd[i]);
          t_filter += std::chrono::duration<double>(tb - ta).count();
          t_tables += std::chrono::duration<double>(clk::now() - tb).count();
          run.candidates(shadow_, i);
          run.mark_prepared(i);
        }
      } catch (...) {
        run.fail(std::current_exception());
      }
    });
  }
  ~Preparer() {
    stop_.store(true);
    join();
  }
  void join() {
    if (th_.joinable()) th_.join();
  }
  double t_filter = 0, t_tables = 0;  // read after the thread ended

 private:
  gem_ctx *c_;
  gem_ctx *shadow_ = nullptr;
  std::atomic<bool> stop_{false};
  std::thread th_;
};

static bool verbose() {
  const char *v = std::getenv("GEM_VERBOSE");
  return v && v[0] == '1';
}

// Timing and pool state shared by the batched entry points.
static void end_run(gem_ctx *c, double sum_filter, double sum_tables) {
  hipStream_t s = c->stream;
  GEM_HIP_CHECK(hipEventRecord(c->ev[7], s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  GEM_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
  c->times.join_ms = ms;
  // wall time of the whole run's filter and table passes (host clock)
  c->times.filter_ms = sum_filter * 1e3;
  c->times.tables_ms = sum_tables * 1e3;
  prof_flush(c);
  // the run's level buffers and tables are gone: empty segments beyond half
  // of the device memory go back to the driver (the rest stays for the next
  // run, whose first allocations would otherwise wait for the driver)
  pool_trim(pool_retain_bytes());
}

void run_match_batch_learned(gem_ctx *c, const gem_query_view *qs, uint32_t nq, gem_mlp *mlp, uint32_t max_orders,
                             const gem_batch_opts *opts, gem_join_result *results, uint64_t *level_rows,
                             uint32_t *chosen) {
  if (!c->g.loaded) throw Error{GEM_E_STATE, "no data graph loaded"};
  gem_batch_opts o{};
  if (opts) o = *opts;
  const uint32_t group = o.group ? o.group : std::max<uint32_t>(nq, 1);
  max_orders = std::max<uint32_t>(1, max_orders);
  const bool streamed = c->stream_cache_bytes != 0;
  double sum_filter = 0, sum_tables = 0;
  GEM_HIP_CHECK(hipEventRecord(c->ev[6], c->stream));
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  for (uint32_t g0 = 0; g0 < nq; g0 += group) {
    const uint32_t gn = std::min(group, nq - g0);
    // Queries the run starts with are prepared here; with continuous
    // admission (at most max_active running) the later ones are prepared on a
    // second stream while the run goes on, and start as they become ready.
    // (Not while profiling: kernel statistics are the batch context's.)
    const uint32_t eager = (!streamed && c->prep_overlap && c->max_active && !c->prof.on)
                               ? std::min<uint32_t>(gn, c->max_active + 16)
                               : gn;
    std::vector<const gem_query_view *> views(gn);
    for (uint32_t i = 0; i < gn; ++i) views[i] = &qs[g0 + i];
    gem_ctx *shadow = eager < gn ? prep_ctx(c) : nullptr;
    std::vector<QueryDev> qd(gn);
    // queries prepared on the second stream go back to its pool: the batch
    // stream's reads of them end first, also when the run throws
    struct SyncFirst {
      hipStream_t s;
      bool on;
      ~SyncFirst() {
        if (on) (void)hipStreamSynchronize(s);
      }
    } sync_first{c->stream, eager < gn};
    std::vector<BatchItem> items(gn);
    for (uint32_t i = 0; i < gn; ++i) items[i].qd = &qd[i];
    ItemRun run(c, items, mlp, max_orders, o);
    double t_filter = 0, t_tables = 0;
    const auto t_prep = clk::now();
    for (uint32_t i = 0; i < eager; ++i) {
      const auto ta = clk::now();
      run_filter(c, qd[i], views[i], 0);
      const auto tb = clk::now();
      run_tables(c, qd[i]);
      if (streamed) spill_tables(c, qd[i], c->stream_slot_bytes);
      t_filter += secs(ta, tb);
      t_tables += secs(tb, clk::now());
      run.candidates(c, i);
    }
    {
      std::unique_ptr<Preparer> prep;
      if (eager < gn) prep = std::make_unique<Preparer>(c, shadow, views, qd, run, eager, gn);
      if (verbose()) {
        std::fprintf(stderr, "[batch] %u queries: %u prepared first (filter %.2f s, edge tables %.2f s)%s\n", gn,
                     eager, t_filter, t_tables, eager < gn ? ", the rest during the run on a second stream" : "");
        std::fflush(stderr);
      }
      run.run(eager, streamed, views, results + g0, level_rows ? level_rows + (uint64_t)g0 * GEM_MAX_QV : nullptr,
              chosen ? chosen + g0 : nullptr);
      if (prep) {
        prep->join();  // done: every query has started, so every one was prepared
        t_filter += prep->t_filter;
        t_tables += prep->t_tables;
      }
    }
    sum_filter += t_filter;
    sum_tables += t_tables;
    if (verbose()) {
      std::fprintf(stderr, "[batch] %u queries done in %.2f s\n", gn, secs(t_prep, clk::now()));
      std::fflush(stderr);
    }
  }
  end_run(c, sum_filter, sum_tables);
}

void run_match_relaxed(gem_ctx *c, const gem_query_view *q, uint32_t threshold,
                       const std::vector<gemh::EdgeSet> &sets, gem_mlp *mlp, uint32_t max_orders,
                       const gem_batch_opts *opts, gem_join_result *results) {
  if (!c->g.loaded) throw Error{GEM_E_STATE, "no data graph loaded"};
  const uint32_t n = (uint32_t)sets.size();
  std::vector<std::vector<uint32_t>> edges(n);
  std::vector<gem_query_view> views(n);
  for (uint32_t i = 0; i < n; ++i) {
    gemh::subquery_edges(q, sets[i], &edges[i]);
    views[i] = gem_query_view{q->num_v, (uint32_t)(edges[i].size() / 2), q->labels, edges[i].data()};
  }
  const bool streamed = c->stream_cache_bytes != 0;
  if (!c->relaxed_share || streamed) {
    // every relaxed query filtered and tabled on its own edges
    run_match_batch_learned(c, views.data(), n, mlp, max_orders, opts, results, nullptr, nullptr);
    return;
  }
  // Shared: one filter of the query at the threshold serves every relaxed
  // query (a relaxed query leaves out at most `threshold` edges, each costing
  // a vertex at most one of degree and of neighbour-label count), and one
  // table per query edge serves every relaxed query that keeps the edge.
  gem_batch_opts o{};
  if (opts) o = *opts;
  max_orders = std::max<uint32_t>(1, max_orders);
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  GEM_HIP_CHECK(hipEventRecord(c->ev[6], c->stream));
  QueryDev parent;
  const auto ta = clk::now();
  run_filter(c, parent, q, threshold);
  const auto tb = clk::now();
  run_tables(c, parent);
  const auto tc = clk::now();
  std::vector<gemh::QueryInfo> Qi(n);
  std::vector<std::vector<uint32_t>> emap(n);
  for (uint32_t i = 0; i < n; ++i) {
    std::string why;
    const int st = gemh::analyse_query(&views[i], &Qi[i], &why);
    if (st) throw Error{st, "relaxed query: " + why};
    for (uint32_t e = 0; e < q->num_e; ++e)
      if (sets[i][e]) emap[i].push_back(e);
  }
  const uint32_t group = o.group ? o.group : std::max<uint32_t>(n, 1);
  for (uint32_t g0 = 0; g0 < n; g0 += group) {
    const uint32_t gn = std::min(group, n - g0);
    std::vector<BatchItem> items(gn);
    for (uint32_t i = 0; i < gn; ++i) items[i] = BatchItem{&parent, &Qi[g0 + i], &emap[g0 + i]};
    std::vector<const gem_query_view *> vv(gn);
    for (uint32_t i = 0; i < gn; ++i) vv[i] = &views[g0 + i];
    ItemRun run(c, items, mlp, max_orders, o);
    for (uint32_t i = 0; i < gn; ++i) run.candidates(c, i);
    run.run(gn, false, vv, results + g0, nullptr, nullptr);
  }
  if (verbose()) {
    std::fprintf(stderr, "[relaxed] %u relaxed queries sharing one filter (%.2f s) and one table build (%.2f s)\n",
                 n, secs(ta, tb), secs(tb, tc));
    std::fflush(stderr);
  }
  end_run(c, secs(ta, tb), secs(tb, tc));
}

void run_vjoin(gem_ctx *c, const uint32_t *vorder, uint64_t cap, gem_join_result *out, uint64_t *level_rows) {
  QueryDev &Qd = c->q;
  if (!Qd.tabled) throw Error{GEM_E_STATE, "edge tables have not been built"};
  VQueryState S;
  plan_query(c, S, Qd, vorder);
  std::vector<VQueryState *> ptrs{&S};
  VBatchRunner runner(c, cap, c->work_limit_rows, c->level_budget_bytes);
  runner.run(ptrs);
  fill_result(S, cap, out, level_rows);
  prof_flush(c);
}

}  // namespace gemd
