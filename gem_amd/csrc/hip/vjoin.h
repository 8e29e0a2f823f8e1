// Vertex-at-a-time batched join: device descriptors and the host driver.
#pragma once

#include <memory>
#include <vector>

#include "../host/vsched.h"
#include "ctx.h"

namespace gemd {

// One level of one query: extend rows (positions of vorder[0..i-1]) by
// vorder[i].  Device copy, one per (query, level).
struct VLevel {
  uint32_t nb, ndup;
  uint32_t key_col[GEM_MAX_QV];
  const uint64_t *toff[GEM_MAX_QV];
  const uint32_t *tval[GEM_MAX_QV];  // in-HBM tables
  uint32_t dup_col[GEM_MAX_QV];
  const uint32_t *dup_can[GEM_MAX_QV];
  const uint64_t *bm_new;
  const uint32_t *rank_new;
  // streamed tables: block map of table j (address of the values of key p =
  // vmap[j][p >> kshift[j]] + 4 * toff[j][p]) and its first global block id
  const uint64_t *vmap[GEM_MAX_QV];
  uint32_t kshift[GEM_MAX_QV];
  uint32_t gblk[GEM_MAX_QV];
};

// One unit of a launch: a row range of one pending level of one query.
struct VItem {
  const VLevel *lv;
  const uint32_t *rows;
  uint64_t row_begin, row_end;
  uint32_t k, pad;
  uint32_t *out;
  uint64_t out_cap;
  unsigned long long *counter;
};

// A row whose driving list is too long for one wave: walked in chunks by
// k_vext_heavy (t = item of the launch, r = row of the item's level).
struct HeavyRow {
  uint32_t t, len;
  uint64_t r;
};

struct VQueryState {
  const QueryDev *qd = nullptr;
  // A relaxed query sharing its parent's candidates and tables: planned on its
  // own edges (Qp), whose edge i is edge (*edge_map)[i] of qd's query; the
  // plan's table ids are qd's.  Null: the query is qd->Q itself.
  const gemh::QueryInfo *Qp = nullptr;
  const std::vector<uint32_t> *edge_map = nullptr;
  gemh::VertexPlan plan;
  DevBuf levels;  // VLevel[nv]
  std::vector<VLevel> hlv;  // host copy of `levels`
  std::vector<uint32_t> max_nb, max_ndup;  // per level
  // depth-first stack of pending level ranges, counts and limits (host/vsched.h)
  gemh::SQuery<DevBuf> s;
  // Restarts (learned batched path): the vertex orders in the order they are
  // tried and their candidate indices; ranked[cur] is the one running.
  // Attempt a runs ranked[a] while a < size, then the order whose ended
  // attempts counted the most embeddings per unit of work (ranked[a % size]
  // while none counted any), and may add
  // first_attempt * growth^(a / width) work (first_attempt == 0: no restarts;
  // attempt max_attempts - 1, when max_attempts != 0, has no limit).
  // mnb_all / mnd_all: the most backward lists and equal-label columns of a
  // level over every ranked order (the launch buffers are sized for them).
  std::vector<std::vector<uint32_t>> ranked;
  std::vector<uint32_t> ranked_idx;
  uint32_t cur = 0;
  uint64_t first_attempt = 0;
  uint32_t growth = 4, width = 1;
  uint32_t max_attempts = 0;
  std::vector<double> rate;  // per ranked order: most embeddings per unit of work in its ended attempts
  uint32_t mnb_all = 1, mnd_all = 0;
  // A racing twin (SQuery::partner set, index past the batch): attempt a runs
  // ranked[(a + shift) % size] while a < size; shift is set when it starts so
  // that its first attempt takes the order after its partner's current one,
  // at its partner's current limit.
  bool twin = false;
  uint32_t shift = 0;
  // Not yet complete when the run starts (ranked orders and plan missing):
  // a query prepared on another stream during the run, or a twin; completed
  // when it starts.  max_nb / max_ndup then hold one upper bound each (a
  // level has at most the degree of its vertex in backward lists and one
  // equal-label column less than the vertices of its label), which sizes the
  // run's launch buffers.
  bool pending = false;
  uint32_t pending_nv = 0;
};

// Completion of pending primaries (see VQueryState::pending): ready() says
// whether query q's filter, tables and candidate orders exist; complete()
// waits for them, then ranks its orders and plans it.
struct LazyStates {
  virtual bool ready(uint32_t q) = 0;
  virtual void complete(uint32_t q) = 0;
  virtual ~LazyStates() = default;
};

void build_vlevels(gem_ctx *c, VQueryState &S);
// The vertex plan of `vorder` for S (S.qd set) and its device level table.
void replan(gem_ctx *c, VQueryState &S, const std::vector<uint32_t> &vorder);
// Work limit of attempt a (0 = none).
uint64_t attempt_limit(const VQueryState &S, uint32_t a);

class VBatchRunner {
 public:
  // streamed: the queries' tables were spilled to host memory (spill_tables)
  // and are read through a device block cache of the context's size
  VBatchRunner(gem_ctx *c, uint64_t cap, uint64_t work_limit, uint64_t budget_bytes, bool streamed = false)
      : c_(c), cap_(cap), work_limit_(work_limit), budget_bytes_(budget_bytes), streamed_(streamed) {}
  // queries [0, n) are the batch, later ones race a partner among them
  // (SQuery::partner, VQueryState::twin); 0 = no racing
  void set_primaries(uint32_t n) { primaries_ = n; }
  void set_lazy(LazyStates *lazy) { lazy_ = lazy; }
  void run(std::vector<VQueryState *> &qs);
  const gemh::SchedStats &stats() const { return stats_; }

 private:
  gem_ctx *c_;
  uint64_t cap_, work_limit_, budget_bytes_;
  bool streamed_;
  uint32_t primaries_ = 0;
  LazyStates *lazy_ = nullptr;
  gemh::SchedStats stats_;
};

}  // namespace gemd
