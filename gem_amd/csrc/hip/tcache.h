// Streamed edge tables for the batched join (out-of-core form).
//
// After a query's tables are built in HBM, spill_tables() moves their values
// to pinned host memory and keeps only the offsets on the device.  Each table
// is cut into key-range blocks (2^kshift keys each) whose values fit one cache
// slot.  During the join, every iteration of the batched runner
//   1. marks, per block, the first and last 256-row task that reads it
//      (k_vmark: atomicMin / atomicMax of the task index),
//   2. cuts its tasks into groups such that the blocks live in two
//      consecutive groups fit the cache together (a block is live from its
//      first to its last use; when the blocks live at a single task do not
//      fit, the iteration is marked and planned again in smaller windows of
//      tasks, down to one task),
//   3. for each group, loads the missing blocks on a second stream while the
//      previous group's kernels run (one kernel reads them from pinned host
//      memory over the host link into their slots), rewrites their entries
//      of the block map (vmap: block -> device address of its values, minus
//      the block's first offset, so that address = vmap[block] + 4 * offset),
//      and lets the group's kernels start once its copies are done.
// Absent blocks map to a zeroed guard slot, so a read of a block that is not
// resident stays inside device memory (and shows up as a wrong count, never
// as a fault).  The cache keeps blocks across iterations (least recently used
// ones are replaced first; blocks needed later in the same iteration last).
// Slots come in two size classes, so that small blocks (most tables of a
// sparse query) do not each pin a large slot.
#pragma once

#include <vector>

#include "ctx.h"
#include "vjoin.h"

namespace gemd {

// Move the values of Qd's tables to pinned host memory (block layout for
// slot_bytes-sized cache slots).  Throws GEM_E_LIMIT if one list is longer
// than a slot.
void spill_tables(gem_ctx *c, QueryDev &Qd, uint64_t slot_bytes);

class TableCache {
 public:
  struct Group {
    uint64_t a = 0, b = 0;  // task range [a, b) of the window
    bool serial = false;    // its copies wait for the previous group's kernels
    size_t p0 = 0, p1 = 0;  // its entries in the patch list
    size_t l0 = 0, l1 = 0;  // its entries in the load list
  };
  static constexpr uint64_t SMALL_SLOT = 64 << 10;  // size class of small blocks

  TableCache(gem_ctx *c, uint64_t cache_bytes, uint64_t slot_bytes);
  ~TableCache();
  TableCache(const TableCache &) = delete;
  TableCache &operator=(const TableCache &) = delete;

  // Assign global block ids to the queries' tables, size the two slot pools,
  // then patch and upload the level descriptors (vmap, kshift, gblk).
  void attach(std::vector<VQueryState *> &qs);
  // Point query i's level descriptors (its current plan) at the block map
  // and upload them (after attach, and after a restart replanned it).
  void bind(VQueryState &S, size_t i);

  uint32_t *d_first() const { return first_.as<uint32_t>(); }
  uint32_t *d_last() const { return last_.as<uint32_t>(); }

  // On the context stream: first = ~0, last = 0 for every block.
  void reset_marks();
  // Read the marks back (synchronises the context stream), then cut
  // total_tasks into groups and decide every copy and eviction.  False (and
  // no plan) when the blocks live at one task cannot all be resident: the
  // caller marks a smaller window of tasks and plans again.
  bool plan(uint64_t total_tasks);
  const std::vector<Group> &groups() const { return groups_; }
  // Device counters for the group launches, zeroed: per group 3 work queues,
  // 2 heavy-row queues and 2 heavy-row counts.
  unsigned int *queues();
  // Copy stream: group g's block loads and map updates; the context stream
  // waits for them.
  void issue(size_t g);
  // Context stream: mark the end of group g's kernels.
  void done(size_t g);
  // After the iteration's stream synchronisation: accumulate statistics.
  void finish();

 private:
  std::vector<std::vector<uint64_t>> base_;  // first global block of each (query, directed table)
  gem_ctx *c_;
  uint64_t cache_bytes_;
  uint64_t sbytes_[2];      // slot bytes of the small and the large class
  uint32_t nslots_[2] = {0, 0};
  uint64_t cbase_[2] = {0, 0}, pool_bytes_ = 0;
  hipStream_t copy_ = nullptr;
  DevBuf pool_;  // small slots | large slots | zeroed guard (one large slot)
  DevBuf vmap_, first_, last_, dpatch_, dload_, queues_;
  PinnedBuf h_first_, h_last_, h_patch_, h_load_;
  uint64_t nblk_ = 0;
  std::vector<uint64_t> bsrc_;          // device-visible address of block values
  std::vector<uint64_t> boff_, blen_;   // first offset and length (entries)
  std::vector<uint8_t> cls_;            // block -> size class
  std::vector<uint32_t> blk_slot_;      // block -> slot (in its class) or NONE
  std::vector<uint64_t> slot_blk_[2];   // slot -> block or NONE
  std::vector<uint64_t> stamp_[2];      // slot -> last use
  std::vector<int64_t> live_in_;        // block -> id of the last group it was live in
  int64_t gbase_ = 0;                   // id of the window's first group
  uint64_t clock_ = 0;
  std::vector<Group> groups_;
  std::vector<hipEvent_t> ev_;          // 3 per group: copy begin, copy end, kernels end
  hipEvent_t ev_step0_ = nullptr;
  uint64_t hits_ = 0;
  void collect();
  uint64_t guard_addr() const;
  uint64_t slot_addr(int cls, uint32_t s) const;
};

}  // namespace gemd
