// Query analysis, GEM ordering and the edge-at-a-time join plan (host C++).
#pragma once

#include <bitset>
#include <cstdint>
#include <string>
#include <vector>

#include "gem_hip.h"

namespace gemh {

using EdgeSet = std::bitset<GEM_MAX_QE>;

struct QueryInfo {
  uint32_t nv = 0, ne = 0;
  std::vector<uint32_t> lab;                          // original labels
  std::vector<std::pair<uint32_t, uint32_t>> edges;   // (src, dst), src < dst
  std::vector<uint32_t> deg;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj;  // (nbr, edge)
};

// Validate and normalise a query view.  Rejects (GEM_E_QUERY) queries that are
// disconnected, have fewer than 2 edges (the reference computes nothing for
// them, in-core/lib/plan/src/relax.cpp:89), self loops, repeated edges, or
// exceed GEM_MAX_QV / GEM_MAX_QE.
int analyse_query(const gem_query_view *q, QueryInfo *out, std::string *why);

// Is the edge-induced subgraph of `s` connected?  (cpuGraph::check_connectivity,
// in-core/lib/graph/src/cpuGraph.cpp:74-194.)
bool edges_connected(const QueryInfo &q, const EdgeSet &s);

// Does the edge set `s` connect all q.nv query vertices?  (checkConnectivity,
// in-core/lib/plan/src/relax.cpp:57-81: one union-find component.)
bool edges_span(const QueryInfo &q, const EdgeSet &s);

// The relaxed queries of q at `threshold` (relax(), in-core/lib/plan/src/
// relax.cpp:84-142): every edge subset that leaves out at most `threshold`
// edges and still connects all vertices, full set first, then by the number
// left out and, within one number, by the ascending tuple of left-out edge
// indices.  GEM_E_LIMIT once more than max_rq exist.
int relaxed_edge_sets(const QueryInfo &q, uint32_t threshold, uint64_t max_rq, std::vector<EdgeSet> *out);

// masks[i * GEM_RQ_WORDS + w] bit b = edge 64 w + b of sets[i].
inline void pack_edge_sets(const std::vector<EdgeSet> &sets, uint64_t *masks) {
  for (size_t i = 0; i < sets.size(); ++i)
    for (uint32_t w = 0; w < GEM_RQ_WORDS; ++w) {
      uint64_t m = 0;
      for (uint32_t b = 0; b < 64; ++b)
        if (sets[i][64 * w + b]) m |= 1ull << b;
      masks[i * GEM_RQ_WORDS + w] = m;
    }
}

// The query view of q restricted to the edges of s (edges in index order).
inline void subquery_edges(const gem_query_view *q, const EdgeSet &s, std::vector<uint32_t> *edges) {
  edges->clear();
  for (uint32_t i = 0; i < q->num_e; ++i)
    if (s[i]) {
      edges->push_back(q->edges[2 * i]);
      edges->push_back(q->edges[2 * i + 1]);
    }
}

// GEM order; see gem_plan_gem_order in include/gem_hip.h.
int gem_order(const QueryInfo &q, const uint64_t *table_sizes, uint32_t *order);

// Is `order` a permutation of the edges whose every prefix is connected?
bool order_is_connected(const QueryInfo &q, const uint32_t *order);

struct JoinStep {
  uint32_t edge = 0;
  bool select = false;
  uint32_t directed = 0;   // directed edge id of the table used (2e or 2e+1)
  uint32_t key_col = 0;    // column whose position keys the table
  uint32_t val_col = 0;    // select: column tested for membership
  uint32_t new_u = 0;      // expand: query vertex of the new column
  uint32_t num_dup = 0;    // expand: columns with the new vertex's label
  uint32_t dup_cols[GEM_MAX_QV] = {};
};

struct JoinPlan {
  uint32_t first_edge = 0;
  uint32_t first_directed = 0;  // table that seeds the rows: a -> b
  std::vector<uint32_t> col_u;  // query vertex of each column, in join order
  std::vector<JoinStep> steps;  // steps after the first edge
};

// Column/step bookkeeping of the reference's process()
// (in-core/lib/join/src/join_trie.cu:1025-1085): the first edge's rows come
// from the table keyed on its smaller endpoint unless that endpoint is a leaf
// (in-core/lib/structure/src/unifiedTrie.cu:1095-1191); an edge whose both
// endpoints are columns is a select keyed on the earlier column; otherwise an
// expansion keyed on the present endpoint, with an injectivity check against
// the columns of equal label (join_trie.cu:450-462).
int make_join_plan(const QueryInfo &q, const uint32_t *order, JoinPlan *plan);

// Vertex-at-a-time plan: level i adds query vertex vorder[i] and applies every
// query edge between it and an earlier vertex (its backward neighbours) at
// once, as an intersection of their tables; equal-label earlier vertices are
// excluded (injectivity).  The final count does not depend on the order (the
// count of injective, edge-preserving maps into the candidate sets).
struct VLevelPlan {
  uint32_t nb = 0;                   // backward neighbours
  uint32_t key_col[GEM_MAX_QV] = {};
  uint32_t directed[GEM_MAX_QV] = {};  // table key -> new vertex
  uint32_t ndup = 0;
  uint32_t dup_col[GEM_MAX_QV] = {};
};
struct VertexPlan {
  std::vector<uint32_t> vorder;       // query vertex of each column
  std::vector<VLevelPlan> levels;     // levels[i] extends to vorder[i] (i >= 1)
};
// Vertex order of first appearance along a connected edge order, starting at
// the non-leaf endpoint of the first edge (as the reference's first table).
int vertex_order_from_edges(const QueryInfo &q, const uint32_t *edge_order, std::vector<uint32_t> *vorder);
// Requires vorder[0] to be a non-leaf and every later vertex to have an
// earlier neighbour.
int make_vertex_plan(const QueryInfo &q, const std::vector<uint32_t> &vorder, VertexPlan *plan);

// ------------------------------------------------------- learned ordering
// Statistics of a filtered query that the order features are built from.
struct QueryStats {
  std::vector<uint64_t> ncan;        // |Can(u)|
  std::vector<uint64_t> tsize;       // |T_d| per directed edge (0 if absent)
  std::vector<uint64_t> label_freq;  // data frequency of u's label
  uint64_t data_v = 0, data_e = 0;
};
constexpr uint32_t FEAT_PER_LEVEL = 8;
constexpr uint32_t FEAT_LEVELS = GEM_MAX_QV;
constexpr uint32_t FEAT_GLOBAL = 16;
constexpr uint32_t FEAT_DIM = FEAT_LEVELS * FEAT_PER_LEVEL + FEAT_GLOBAL;  // 272
// int32 features of one vertex order (log2-quantised counts, x16).
void order_features(const QueryInfo &q, const QueryStats &s, const std::vector<uint32_t> &vorder, int32_t *feat);
// Up to `max_orders` distinct valid vertex orders: the GEM-derived order
// first, then greedy orders by the independence estimate from several
// starts, then seeded random connected orders.
void candidate_orders(const QueryInfo &q, const QueryStats &s, const std::vector<uint32_t> &gem_vorder,
                      uint32_t max_orders, uint64_t seed, std::vector<std::vector<uint32_t>> *out);

}  // namespace gemh
