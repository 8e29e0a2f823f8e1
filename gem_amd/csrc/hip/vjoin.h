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
  gemh::VertexPlan plan;
  DevBuf levels;  // VLevel[nv]
  std::vector<VLevel> hlv;  // host copy of `levels`
  std::vector<uint32_t> max_nb, max_ndup;  // per level
  // depth-first stack of pending level ranges, counts and limits (host/vsched.h)
  gemh::SQuery<DevBuf> s;
};

void build_vlevels(gem_ctx *c, VQueryState &S);

class VBatchRunner {
 public:
  // streamed: the queries' tables were spilled to host memory (spill_tables)
  // and are read through a device block cache of the context's size
  VBatchRunner(gem_ctx *c, uint64_t cap, uint64_t work_limit, uint64_t budget_bytes, bool streamed = false)
      : c_(c), cap_(cap), work_limit_(work_limit), budget_bytes_(budget_bytes), streamed_(streamed) {}
  void run(std::vector<VQueryState *> &qs);
  const gemh::SchedStats &stats() const { return stats_; }

 private:
  gem_ctx *c_;
  uint64_t cap_, work_limit_, budget_bytes_;
  bool streamed_;
  gemh::SchedStats stats_;
};

}  // namespace gemd
