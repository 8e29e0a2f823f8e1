// GEM ordering at threshold 0 without the edge-subset lattice.
//
// The reference scores candidates in find_best_edge_candidate
// (in-core/lib/structure/src/unifiedTrie.cu:939-1018) over the lattice of
// connected edge subsets built by Lattice::construct_linked_list_extend
// (in-core/lib/structure/src/lattice.cpp:41-185).  At threshold 0 the only
// relaxed query is the full query, so for every single edge the contribution
// value (lattice.cpp:443-467) is c = 1 / (1 + 0.1 (|E|-1)), and a single-edge
// node's lattice parents are exactly its connected 2-edge supersets.
//
// What remains of the lattice is the order in which the reference scans the
// candidates of a "next edge" step: ascending lattice id.  Ids are handed out
// breadth-first from the full edge set, a node getting its id from the first
// (smallest-id) parent that reaches it, parents expanding their removable
// edges in ascending index.  By induction on the level, a node's id order
// within its level is the lexicographic order of its lexicographically
// smallest removal sequence (remove edges of E \ N one at a time, keeping the
// remainder connected), and that sequence is produced greedily: always remove
// the smallest-index edge whose removal keeps the remainder connected (some
// such edge exists while the remainder strictly contains a connected N).  So
// ties are broken by comparing greedy sequences; no lattice is built, and the
// order is defined for queries far beyond the reference's 32-edge bitset.
#include "planner.h"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "host_graph.h"

namespace gemh {

int analyse_query(const gem_query_view *q, QueryInfo *out, std::string *why) {
  auto fail = [&](int code, const std::string &m) {
    if (why) *why = m;
    return code;
  };
  if (!q || !out || !q->labels || (q->num_e && !q->edges)) return fail(GEM_E_INVALID, "null query");
  if (q->num_v > GEM_MAX_QV) return fail(GEM_E_LIMIT, "query has more than GEM_MAX_QV vertices");
  if (q->num_e > GEM_MAX_QE) return fail(GEM_E_LIMIT, "query has more than GEM_MAX_QE edges");
  if (q->num_e < 2) return fail(GEM_E_QUERY, "query needs at least 2 edges");
  QueryInfo &Q = *out;
  Q.nv = q->num_v;
  Q.ne = q->num_e;
  Q.lab.assign(q->labels, q->labels + q->num_v);
  Q.edges.resize(Q.ne);
  Q.deg.assign(Q.nv, 0);
  Q.adj.assign(Q.nv, {});
  std::vector<uint64_t> keys;
  for (uint32_t i = 0; i < Q.ne; ++i) {
    uint32_t s = q->edges[2 * i], d = q->edges[2 * i + 1];
    if (s >= Q.nv || d >= Q.nv) return fail(GEM_E_INVALID, "edge endpoint out of range");
    if (s == d) return fail(GEM_E_QUERY, "self loop in query");
    if (s > d) std::swap(s, d);
    Q.edges[i] = {s, d};
    Q.deg[s]++;
    Q.deg[d]++;
    Q.adj[s].push_back({d, i});
    Q.adj[d].push_back({s, i});
    keys.push_back(((uint64_t)s << 32) | d);
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) return fail(GEM_E_QUERY, "repeated query edge");
  EdgeSet all;
  for (uint32_t i = 0; i < Q.ne; ++i) all.set(i);
  for (uint32_t v = 0; v < Q.nv; ++v)
    if (Q.deg[v] == 0) return fail(GEM_E_QUERY, "isolated query vertex");
  if (!edges_connected(Q, all)) return fail(GEM_E_QUERY, "query is disconnected");
  return GEM_OK;
}

bool edges_connected(const QueryInfo &q, const EdgeSet &s) {
  uint32_t parent[GEM_MAX_QV];
  for (uint32_t v = 0; v < q.nv; ++v) parent[v] = v;
  auto find = [&](uint32_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  int first = -1;
  for (uint32_t i = 0; i < q.ne; ++i) {
    if (!s[i]) continue;
    if (first < 0) first = (int)i;
    uint32_t a = find(q.edges[i].first), b = find(q.edges[i].second);
    if (a != b) parent[a] = b;
  }
  if (first < 0) return false;
  uint32_t r = find(q.edges[first].first);
  for (uint32_t i = 0; i < q.ne; ++i)
    if (s[i] && find(q.edges[i].first) != r) return false;
  return true;
}

bool edges_span(const QueryInfo &q, const EdgeSet &s) {
  uint32_t parent[GEM_MAX_QV];
  for (uint32_t v = 0; v < q.nv; ++v) parent[v] = v;
  auto find = [&](uint32_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  uint32_t components = q.nv;
  for (uint32_t i = 0; i < q.ne; ++i) {
    if (!s[i]) continue;
    uint32_t a = find(q.edges[i].first), b = find(q.edges[i].second);
    if (a != b) {
      parent[a] = b;
      --components;
    }
  }
  return components <= 1;
}

// The reference reaches its edge sets breadth-first, one removal at a time,
// keeping a set only while it spans; a superset of a spanning set spans, so
// that reaches exactly the spanning sets with at most `threshold` edges left
// out.  Here they are listed depth-first over ascending left-out tuples, a
// branch ending as soon as its set stops spanning (every extension of it is a
// subset and does not span either), then ordered by size.
int relaxed_edge_sets(const QueryInfo &q, uint32_t threshold, uint64_t max_rq, std::vector<EdgeSet> *out) {
  std::vector<std::pair<uint32_t, std::vector<uint32_t>>> found;  // (left out, tuple)
  EdgeSet cur;
  for (uint32_t i = 0; i < q.ne; ++i) cur.set(i);
  std::vector<uint32_t> tuple;
  bool over = false;
  auto dfs = [&](auto &&self, uint32_t from) -> void {
    if (over) return;
    if (found.size() >= max_rq) {
      over = true;
      return;
    }
    found.push_back({(uint32_t)tuple.size(), tuple});
    if (tuple.size() >= threshold) return;
    for (uint32_t e = from; e < q.ne && !over; ++e) {
      cur.reset(e);
      if (edges_span(q, cur)) {
        tuple.push_back(e);
        self(self, e + 1);
        tuple.pop_back();
      }
      cur.set(e);
    }
  };
  dfs(dfs, 0);
  if (over) return GEM_E_LIMIT;
  std::stable_sort(found.begin(), found.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  out->clear();
  out->reserve(found.size());
  for (auto &f : found) {
    EdgeSet s;
    for (uint32_t i = 0; i < q.ne; ++i) s.set(i);
    for (uint32_t e : f.second) s.reset(e);
    out->push_back(s);
  }
  return GEM_OK;
}

// Greedy (lexicographically smallest) removal sequence from the full edge set
// down to `target`.
static std::vector<uint32_t> removal_sequence(const QueryInfo &q, const EdgeSet &target) {
  EdgeSet cur;
  for (uint32_t i = 0; i < q.ne; ++i) cur.set(i);
  std::vector<uint32_t> seq;
  while (cur != target) {
    bool removed = false;
    for (uint32_t f = 0; f < q.ne; ++f) {
      if (!cur[f] || target[f]) continue;
      EdgeSet t = cur;
      t.reset(f);
      if (edges_connected(q, t)) {
        seq.push_back(f);
        cur = t;
        removed = true;
        break;
      }
    }
    if (!removed) break;  // unreachable for a connected target
  }
  return seq;
}

static bool touches(const QueryInfo &q, const EdgeSet &cur, uint32_t e) {
  // edge e shares a vertex with the edge-induced subgraph of cur
  uint32_t a = q.edges[e].first, b = q.edges[e].second;
  for (auto [n, f] : q.adj[a])
    if (cur[f]) return true;
  for (auto [n, f] : q.adj[b])
    if (cur[f]) return true;
  (void)a;
  return false;
}

int gem_order(const QueryInfo &q, const uint64_t *sz, uint32_t *order) {
  // lattice.cpp:457-463: decay of the single full-query target at distance |E|-1
  const double contribution = 0.0 + 1.0 / (1.0 + 0.1 * (int)(q.ne - 1));
  // First edge (unifiedTrie.cu:946-973): strict '>' over edge index order.
  double best = -1.0;
  uint32_t first = UINT32_MAX;
  for (uint32_t i = 0; i < q.ne; ++i) {
    const double size_factor = 1.0 / (1.0 + (double)sz[i] + 1);
    const uint64_t supersets = (uint64_t)(q.deg[q.edges[i].first] - 1) + (q.deg[q.edges[i].second] - 1);
    const double connectivity_factor = (double)supersets * 0.1;
    const double score = 0.6 * size_factor + 0.3 * contribution + 0.1 * connectivity_factor;
    if (score > best) {
      best = score;
      first = i;
    }
  }
  EdgeSet cur;
  cur.set(first);
  order[0] = first;
  for (uint32_t step = 1; step < q.ne; ++step) {
    // Next edge (unifiedTrie.cu:978-1016): candidates are the lattice parents of
    // the current node, scanned in ascending lattice id with strict '>'.
    std::vector<uint32_t> cands;
    std::vector<double> scores;
    double mx = -1.0;
    for (uint32_t e = 0; e < q.ne; ++e) {
      if (cur[e] || !touches(q, cur, e)) continue;
      const double size_factor = 1.0 / (1.0 + (double)sz[e] + 1);
      const double score = 0.4 * size_factor + 0.4 * contribution;
      cands.push_back(e);
      scores.push_back(score);
      mx = std::max(mx, score);
    }
    if (cands.empty()) return GEM_E_QUERY;
    uint32_t pick = UINT32_MAX;
    std::vector<uint32_t> pick_seq;
    for (size_t i = 0; i < cands.size(); ++i) {
      if (scores[i] != mx) continue;
      EdgeSet next = cur;
      next.set(cands[i]);
      std::vector<uint32_t> seq = removal_sequence(q, next);
      if (pick == UINT32_MAX || seq < pick_seq) {
        pick = cands[i];
        pick_seq.swap(seq);
      }
    }
    order[step] = pick;
    cur.set(pick);
  }
  return GEM_OK;
}

bool order_is_connected(const QueryInfo &q, const uint32_t *order) {
  EdgeSet cur;
  for (uint32_t i = 0; i < q.ne; ++i) {
    uint32_t e = order[i];
    if (e >= q.ne || cur[e]) return false;
    if (i > 0 && !touches(q, cur, e)) return false;
    cur.set(e);
  }
  return true;
}

int make_join_plan(const QueryInfo &q, const uint32_t *order, JoinPlan *plan) {
  if (!order_is_connected(q, order)) return GEM_E_INVALID;
  JoinPlan &P = *plan;
  std::vector<int> col_of(q.nv, -1);
  const uint32_t e0 = order[0];
  const uint32_t s = q.edges[e0].first, d = q.edges[e0].second;
  const bool key_small = q.deg[s] != 1;
  const uint32_t a = key_small ? s : d, b = key_small ? d : s;
  P.first_edge = e0;
  P.first_directed = 2 * e0 + (key_small ? 0 : 1);
  P.col_u = {a, b};
  col_of[a] = 0;
  col_of[b] = 1;
  P.steps.clear();
  for (uint32_t i = 1; i < q.ne; ++i) {
    const uint32_t e = order[i];
    const uint32_t x = q.edges[e].first, y = q.edges[e].second;  // x < y
    JoinStep st;
    st.edge = e;
    if (col_of[x] >= 0 && col_of[y] >= 0) {
      st.select = true;
      st.key_col = (uint32_t)std::min(col_of[x], col_of[y]);
      st.val_col = (uint32_t)std::max(col_of[x], col_of[y]);
      st.directed = 2 * e + (P.col_u[st.key_col] == x ? 0 : 1);
    } else {
      const bool fwd = col_of[x] >= 0;  // key on the present endpoint
      const uint32_t present = fwd ? x : y, fresh = fwd ? y : x;
      st.select = false;
      st.key_col = (uint32_t)col_of[present];
      st.new_u = fresh;
      st.directed = 2 * e + (fwd ? 0 : 1);
      for (uint32_t c = 0; c < P.col_u.size(); ++c)
        if (q.lab[P.col_u[c]] == q.lab[fresh]) st.dup_cols[st.num_dup++] = c;
      col_of[fresh] = (int)P.col_u.size();
      P.col_u.push_back(fresh);
    }
    P.steps.push_back(st);
  }
  return GEM_OK;
}

int vertex_order_from_edges(const QueryInfo &q, const uint32_t *edge_order, std::vector<uint32_t> *vorder) {
  if (!order_is_connected(q, edge_order)) return GEM_E_INVALID;
  std::vector<char> seen(q.nv, 0);
  vorder->clear();
  const uint32_t s = q.edges[edge_order[0]].first, d = q.edges[edge_order[0]].second;
  const uint32_t a = q.deg[s] != 1 ? s : d, b = a == s ? d : s;
  for (uint32_t v : {a, b}) {
    vorder->push_back(v);
    seen[v] = 1;
  }
  for (uint32_t i = 1; i < q.ne; ++i) {
    const uint32_t x = q.edges[edge_order[i]].first, y = q.edges[edge_order[i]].second;
    for (uint32_t v : {x, y})
      if (!seen[v]) {
        seen[v] = 1;
        vorder->push_back(v);
      }
  }
  return vorder->size() == q.nv ? GEM_OK : GEM_E_INVALID;
}

int make_vertex_plan(const QueryInfo &q, const std::vector<uint32_t> &vorder, VertexPlan *plan) {
  if (vorder.size() != q.nv) return GEM_E_INVALID;
  std::vector<int> col(q.nv, -1);
  for (uint32_t i = 0; i < q.nv; ++i) {
    if (vorder[i] >= q.nv || col[vorder[i]] >= 0) return GEM_E_INVALID;
    col[vorder[i]] = (int)i;
  }
  if (q.deg[vorder[0]] == 1) return GEM_E_INVALID;  // no table is keyed on a leaf
  plan->vorder = vorder;
  plan->levels.assign(q.nv, VLevelPlan());
  for (uint32_t i = 1; i < q.nv; ++i) {
    const uint32_t u = vorder[i];
    VLevelPlan &L = plan->levels[i];
    for (auto [w, e] : q.adj[u]) {
      if (col[w] >= (int)i) continue;
      // table keyed on the earlier vertex w towards u
      L.key_col[L.nb] = (uint32_t)col[w];
      L.directed[L.nb] = 2 * e + (q.edges[e].first == w ? 0 : 1);
      L.nb++;
    }
    if (L.nb == 0) return GEM_E_INVALID;
    for (uint32_t c = 0; c < i; ++c)
      if (q.lab[vorder[c]] == q.lab[u]) L.dup_col[L.ndup++] = c;
  }
  return GEM_OK;
}

static int32_t q16(double x) { return (int32_t)std::llround(16.0 * std::log2(1.0 + std::max(0.0, x))); }

// fanout of the table w -> u: average row length
static double fanout(const QueryInfo &q, const QueryStats &s, uint32_t w, uint32_t u) {
  for (auto [n, e] : q.adj[w])
    if (n == u) {
      const uint32_t d = 2 * e + (q.edges[e].first == w ? 0 : 1);
      uint64_t t = s.tsize[d];
      if (t == 0) t = s.tsize[d ^ 1];  // leaf-keyed direction: same total
      return (double)t / (double)std::max<uint64_t>(1, s.ncan[w]);
    }
  return 0.0;
}

void order_features(const QueryInfo &q, const QueryStats &s, const std::vector<uint32_t> &vo, int32_t *feat) {
  std::fill(feat, feat + FEAT_DIM, 0);
  std::vector<int> col(q.nv, -1);
  for (uint32_t i = 0; i < vo.size(); ++i) col[vo[i]] = (int)i;
  double est = (double)s.ncan[vo[0]], est_sum = 0;
  for (uint32_t i = 0; i < vo.size() && i < FEAT_LEVELS; ++i) {
    const uint32_t u = vo[i];
    int32_t *f = feat + i * FEAT_PER_LEVEL;
    uint32_t nb = 0, ndup = 0;
    double fmin = 1e300, fmax = 0, sel = 1.0;
    std::vector<double> fans;
    for (auto [w, e] : q.adj[u])
      if (col[w] >= 0 && col[w] < (int)i) {
        fans.push_back(fanout(q, s, w, u));
        ++nb;
      }
    for (uint32_t c = 0; c < i; ++c)
      if (q.lab[vo[c]] == q.lab[u]) ++ndup;
    if (nb) {
      std::sort(fans.begin(), fans.end());
      fmin = fans.front();
      fmax = fans.back();
      for (size_t j = 1; j < fans.size(); ++j) sel *= std::min(1.0, fans[j] / std::max<double>(1, s.ncan[u]));
      est = est * fmin * sel;
    } else {
      fmin = 0;
    }
    est_sum += est;
    f[0] = q16((double)s.ncan[u]);
    f[1] = (int32_t)(16 * nb);
    f[2] = (int32_t)(16 * ndup);
    f[3] = q16(fmin);
    f[4] = q16(fmax);
    f[5] = q16(est);
    f[6] = (int32_t)(16 * q.deg[u]);
    f[7] = q16((double)s.label_freq[u]);
  }
  int32_t *g = feat + FEAT_LEVELS * FEAT_PER_LEVEL;
  uint64_t sum_can = 0, max_can = 0, sum_t = 0;
  for (uint32_t u = 0; u < q.nv; ++u) {
    sum_can += s.ncan[u];
    max_can = std::max(max_can, s.ncan[u]);
  }
  for (uint64_t t : s.tsize) sum_t += t;
  std::vector<uint32_t> labs(q.lab);
  std::sort(labs.begin(), labs.end());
  const uint32_t nlab = (uint32_t)(std::unique(labs.begin(), labs.end()) - labs.begin());
  g[0] = (int32_t)(16 * q.nv);
  g[1] = (int32_t)(16 * q.ne);
  g[2] = (int32_t)(16 * nlab);
  g[3] = q16((double)sum_can);
  g[4] = q16((double)max_can);
  g[5] = q16((double)sum_t);
  g[6] = q16(est);
  g[7] = q16(est_sum);
  g[8] = (2 * q.ne >= 3 * q.nv) ? 16 : 0;
  g[9] = q16((double)s.data_v);
  g[10] = q16((double)s.data_e);
}

static bool valid_start(const QueryInfo &q, uint32_t u) { return q.deg[u] != 1; }

void candidate_orders(const QueryInfo &q, const QueryStats &s, const std::vector<uint32_t> &gem_vo,
                      uint32_t max_orders, uint64_t seed, std::vector<std::vector<uint32_t>> *out) {
  out->clear();
  auto add = [&](const std::vector<uint32_t> &vo) {
    if (out->size() >= max_orders) return;
    for (auto &o : *out)
      if (o == vo) return;
    out->push_back(vo);
  };
  // The same order with the query's leaves moved to the end, the leaf with
  // the largest fan-out from its neighbour last.  Valid for any valid order (a
  // leaf is never a non-leaf's only earlier neighbour); a leaf constrains no
  // later vertex, and a last level with a single list is counted without a
  // walk, so saturating queries reach the cap with few rows.
  auto leaves_last = [&](const std::vector<uint32_t> &vo) {
    std::vector<uint32_t> head, leaves;
    for (uint32_t u : vo) (q.deg[u] == 1 && u != vo[0] ? leaves : head).push_back(u);
    std::stable_sort(leaves.begin(), leaves.end(), [&](uint32_t a, uint32_t b) {
      return fanout(q, s, q.adj[a][0].first, a) < fanout(q, s, q.adj[b][0].first, b);
    });
    head.insert(head.end(), leaves.begin(), leaves.end());
    return head;
  };
  std::vector<std::vector<uint32_t>> greedy;
  // greedy by the running estimate: prefer vertices closing more edges, then
  // the smallest estimated growth
  std::vector<uint32_t> starts;
  for (uint32_t u = 0; u < q.nv; ++u)
    if (valid_start(q, u)) starts.push_back(u);
  std::sort(starts.begin(), starts.end(), [&](uint32_t a, uint32_t b) {
    return s.ncan[a] != s.ncan[b] ? s.ncan[a] < s.ncan[b] : a < b;
  });
  for (uint32_t si = 0; si < starts.size() && si < 4; ++si) {
    std::vector<uint32_t> vo{starts[si]};
    std::vector<char> in(q.nv, 0);
    in[starts[si]] = 1;
    while (vo.size() < q.nv) {
      double best = 1e300;
      uint32_t pick = UINT32_MAX, pick_nb = 0;
      for (uint32_t u = 0; u < q.nv; ++u) {
        if (in[u]) continue;
        uint32_t nb = 0;
        double fmin = 1e300, sel = 1.0;
        std::vector<double> fans;
        for (auto [w, e] : q.adj[u])
          if (in[w]) fans.push_back(fanout(q, s, w, u));
        nb = (uint32_t)fans.size();
        if (!nb) continue;
        std::sort(fans.begin(), fans.end());
        fmin = fans.front();
        for (size_t j = 1; j < fans.size(); ++j) sel *= std::min(1.0, fans[j] / std::max<double>(1, s.ncan[u]));
        const double growth = fmin * sel;
        if (nb > pick_nb || (nb == pick_nb && growth < best)) {
          best = growth;
          pick = u;
          pick_nb = nb;
        }
      }
      vo.push_back(pick);
      in[pick] = 1;
    }
    greedy.push_back(vo);
    // a second variant: pure smallest-growth (ignoring closing-edge count)
    std::vector<uint32_t> vo2{starts[si]};
    std::fill(in.begin(), in.end(), 0);
    in[starts[si]] = 1;
    while (vo2.size() < q.nv) {
      double best = 1e300;
      uint32_t pick = UINT32_MAX;
      for (uint32_t u = 0; u < q.nv; ++u) {
        if (in[u]) continue;
        std::vector<double> fans;
        for (auto [w, e] : q.adj[u])
          if (in[w]) fans.push_back(fanout(q, s, w, u));
        if (fans.empty()) continue;
        std::sort(fans.begin(), fans.end());
        double g = fans.front();
        for (size_t j = 1; j < fans.size(); ++j) g *= std::min(1.0, fans[j] / std::max<double>(1, s.ncan[u]));
        if (g < best) {
          best = g;
          pick = u;
        }
      }
      vo2.push_back(pick);
      in[pick] = 1;
    }
    greedy.push_back(vo2);
  }
  // GEM's order first (the fallback), then the greedy orders, each preceded
  // by its leaves-last form
  add(gem_vo);
  if (!greedy.empty()) {
    add(leaves_last(greedy[0]));
    add(greedy[0]);
  }
  add(leaves_last(gem_vo));
  for (size_t i = 1; i < greedy.size(); ++i) {
    add(leaves_last(greedy[i]));
    add(greedy[i]);
  }
  // seeded random connected orders
  Rng rng(seed);
  for (uint32_t tries = 0; out->size() < max_orders && tries < 8 * max_orders; ++tries) {
    std::vector<uint32_t> vo{starts[rng.below(starts.size())]};
    std::vector<char> in(q.nv, 0);
    in[vo[0]] = 1;
    while (vo.size() < q.nv) {
      std::vector<uint32_t> fr;
      for (uint32_t u = 0; u < q.nv; ++u) {
        if (in[u]) continue;
        for (auto [w, e] : q.adj[u])
          if (in[w]) {
            fr.push_back(u);
            break;
          }
      }
      const uint32_t u = fr[rng.below(fr.size())];
      vo.push_back(u);
      in[u] = 1;
    }
    add(vo);
  }
}

}  // namespace gemh

extern "C" int gem_plan_gem_order(const gem_query_view *q, const uint64_t *table_sizes, uint32_t *order_out) {
  if (!table_sizes || !order_out) return GEM_E_INVALID;
  gemh::QueryInfo Q;
  int st = gemh::analyse_query(q, &Q, nullptr);
  if (st) return st;
  return gemh::gem_order(Q, table_sizes, order_out);
}

extern "C" int gem_plan_relaxed_queries(const gem_query_view *q, uint32_t threshold, uint32_t max_rq,
                                        uint64_t *masks, uint32_t *num_rq) {
  if (!num_rq || (max_rq && !masks)) return GEM_E_INVALID;
  gemh::QueryInfo Q;
  int st = gemh::analyse_query(q, &Q, nullptr);
  if (st) return st;
  std::vector<gemh::EdgeSet> sets;
  st = gemh::relaxed_edge_sets(Q, threshold, max_rq ? max_rq : GEM_MAX_RQ, &sets);
  if (st) return st;
  *num_rq = (uint32_t)sets.size();
  if (max_rq) gemh::pack_edge_sets(sets, masks);
  return GEM_OK;
}
