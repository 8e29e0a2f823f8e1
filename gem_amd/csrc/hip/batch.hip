// Host orchestration of the batched path: per group of queries, filter and
// edge tables for each query, a vertex plan per query (given, or derived from
// GEM's edge order), then one VBatchRunner over the whole group.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

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
    auto it = c->g.label_id.find(Q.lab[u]);
    if (it != c->g.label_id.end()) s.label_freq[u] = c->g.label_freq[it->second];
  }
  s.data_v = c->g.n;
  s.data_e = c->g.slots / 2;
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

// Candidate orders, model ranking and one VBatchRunner over `items`; results
// (and level_rows, chosen) indexed like items.  Returns the seconds spent on
// candidate orders and features.
static double run_items(gem_ctx *c, const std::vector<BatchItem> &items, gem_mlp *mlp, uint32_t max_orders,
                        const gem_batch_opts &o, bool streamed, gem_join_result *results, uint64_t *level_rows,
                        uint32_t *chosen) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const size_t n = items.size();
  std::vector<VQueryState> st(n);
  std::vector<VQueryState *> ptrs;
  std::vector<std::vector<std::vector<uint32_t>>> cand(n);
  std::vector<int32_t> feats;
  std::vector<uint32_t> owner;
  for (size_t i = 0; i < n; ++i) {
    const gemh::QueryInfo &Q = item_query(items[i]);
    const gemh::QueryStats qsx = item_stats(c, items[i]);
    // without a model, candidates beyond GEM's order matter only to restarts
    const uint32_t norders = (mlp || o.first_attempt) ? max_orders : 1;
    gemh::candidate_orders(Q, qsx, item_gem_vertex_order(items[i]), norders, query_seed(Q), &cand[i]);
    if (mlp) {
      for (auto &vo : cand[i]) {
        const size_t off = feats.size();
        feats.resize(off + gemh::FEAT_DIM);
        gemh::order_features(Q, qsx, vo, feats.data() + off);
        owner.push_back((uint32_t)i);
      }
    }
  }
  // Ranking of each query's candidates: without a model, as generated (GEM's
  // order first); with one, by predicted cost, except that GEM's order
  // (candidate 0) stays first unless the cheapest prediction beats it by
  // more than the model's margin.  The first ranked order runs; restarts
  // take the next ones.
  std::vector<std::vector<uint32_t>> rank(n);
  for (size_t i = 0; i < n; ++i) {
    rank[i].resize(cand[i].size());
    for (uint32_t j = 0; j < cand[i].size(); ++j) rank[i][j] = j;
  }
  if (mlp && !owner.empty()) {
    std::vector<float> pred(owner.size());
    const int rc = gem_mlp_predict(mlp, feats.data(), (uint32_t)owner.size(), pred.data());
    if (rc) throw Error{rc, "order model: " + c->err};
    const float margin = mlp_margin(mlp);
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
      const float *p = pred.data() + k;
      k += cand[i].size();
      std::vector<uint32_t> &R = rank[i];
      std::stable_sort(R.begin(), R.end(), [&](uint32_t a, uint32_t b) { return p[a] < p[b]; });
      if (!(p[R[0]] < p[0] - margin)) {
        R.erase(std::find(R.begin(), R.end(), 0u));
        R.insert(R.begin(), 0u);
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    VQueryState &V = st[i];
    V.qd = items[i].qd;
    V.Qp = items[i].Qp;
    V.edge_map = items[i].edge_map;
    V.first_attempt = o.first_attempt;
    V.growth = o.restart_growth ? o.restart_growth : 4;
    V.width = o.restart_width ? o.restart_width : 1;
    V.max_attempts = o.max_attempts;
    for (uint32_t j : rank[i]) {
      V.ranked.push_back(cand[i][j]);
      V.ranked_idx.push_back(j);
    }
    // size the launch buffers for every order a restart may run
    if (attempt_limit(V, 0))
      for (size_t j = 1; j < V.ranked.size(); ++j) {
        gemh::VertexPlan P;
        if (gemh::make_vertex_plan(item_query(items[i]), V.ranked[j], &P))
          throw Error{GEM_E_STATE, "invalid candidate vertex order"};
        for (size_t l = 1; l < P.levels.size(); ++l) {
          V.mnb_all = std::max(V.mnb_all, P.levels[l].nb);
          V.mnd_all = std::max(V.mnd_all, P.levels[l].ndup);
        }
      }
    replan(c, V, V.ranked[0]);
    V.cur = 0;
    ptrs.push_back(&V);
  }
  // Racing: each query with restarts gets a twin on the same candidates and
  // tables; the twins start, in the same order, once every query of the batch
  // has started, when a running query ends and leaves room (the tail, where
  // the few queries left would leave most of the device idle), and whichever
  // of a pair ends first with a count cancels the other.  Counts do not
  // depend on the order, so either gives the query's result.
  std::vector<VQueryState> twins;
  std::vector<int64_t> twin_of(n, -1);
  if (c->racing && !streamed && o.first_attempt) {
    twins.resize(n);
    for (size_t i = 0; i < n; ++i) {
      VQueryState &P = st[i];
      if (P.ranked.size() < 2) continue;
      VQueryState &T = twins[i];
      T.qd = P.qd;
      T.Qp = P.Qp;
      T.edge_map = P.edge_map;
      T.first_attempt = P.first_attempt;
      T.growth = P.growth;
      T.width = P.width;
      T.max_attempts = P.max_attempts;
      T.ranked = P.ranked;
      T.ranked_idx = P.ranked_idx;
      T.mnb_all = P.mnb_all;
      T.mnd_all = P.mnd_all;
      T.twin = true;
      replan(c, T, T.ranked[1]);
      twin_of[i] = (int64_t)ptrs.size();
      T.s.partner = (int64_t)i;
      P.s.partner = twin_of[i];
      ptrs.push_back(&T);
    }
  }
  const double t_orders = std::chrono::duration<double>(clk::now() - t0).count();
  VBatchRunner runner(c, o.cap, o.work_limit, o.level_budget, streamed);
  if (!twins.empty()) runner.set_primaries((uint32_t)n);
  runner.run(ptrs);
  for (size_t i = 0; i < n; ++i) {
    const VQueryState *W = &st[i];
    uint64_t twin_work = 0;
    if (twin_of[i] >= 0) {
      const VQueryState &T = twins[i];
      twin_work = T.s.work;
      // the twin's result when it ended with a count and the query did not
      if (T.s.done && !T.s.cancelled && !T.s.work_limited && (st[i].s.cancelled || st[i].s.work_limited)) W = &T;
    }
    fill_result(*W, o.cap, &results[i], level_rows ? level_rows + i * GEM_MAX_QV : nullptr);
    results[i].work = st[i].s.work + twin_work;
    if (chosen) chosen[i] = W->ranked_idx[W->cur];
  }
  return t_orders;
}

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
    std::vector<QueryDev> qd(gn);
    std::vector<BatchItem> items(gn);
    double t_filter = 0, t_tables = 0;
    const auto t_prep = clk::now();
    for (uint32_t i = 0; i < gn; ++i) {
      const auto ta = clk::now();
      run_filter(c, qd[i], &qs[g0 + i], 0);
      const auto tb = clk::now();
      run_tables(c, qd[i]);
      if (streamed) spill_tables(c, qd[i], c->stream_slot_bytes);
      t_filter += secs(ta, tb);
      t_tables += secs(tb, clk::now());
      items[i].qd = &qd[i];
    }
    if (verbose()) {
      std::fprintf(stderr, "[batch] %u queries: filter %.2f s, edge tables %.2f s\n", gn, t_filter, t_tables);
      std::fflush(stderr);
    }
    sum_filter += t_filter;
    sum_tables += t_tables;
    const double t_orders = run_items(c, items, mlp, max_orders, o, streamed, results + g0,
                                      level_rows ? level_rows + (uint64_t)g0 * GEM_MAX_QV : nullptr,
                                      chosen ? chosen + g0 : nullptr);
    if (verbose()) {
      std::fprintf(stderr, "[batch] %u queries done in %.2f s (candidate orders, features and plans %.2f s)\n", gn,
                   secs(t_prep, clk::now()), t_orders);
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
    run_items(c, items, mlp, max_orders, o, false, results + g0, nullptr, nullptr);
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
