// Streamed edge tables: spill to pinned host memory, block cache in HBM,
// per-iteration copy plan.  The join kernels that read through the block map
// and the marking kernel are in vjoin.hip.
#include <algorithm>
#include <cstring>

#include "tcache.h"

namespace gemd {

namespace {

constexpr uint32_t NONE32 = 0xFFFFFFFFu;
constexpr uint64_t NONE64 = ~0ull;

struct Patch {
  uint64_t idx, val;
};

// One block to bring in: `words` u32 values from src (device-visible address
// of pinned host memory) to dst (a cache slot).
struct Load {
  uint64_t src, dst, words;
};

// A workgroup per block: consecutive lanes read consecutive words, so each
// wave instruction reads 256 contiguous bytes of host memory over the link.
__global__ void __launch_bounds__(256) k_load_blocks(const Load *__restrict__ ld, uint64_t n) {
  for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const Load L = ld[i];
    const uint32_t *src = reinterpret_cast<const uint32_t *>(L.src);
    uint32_t *dst = reinterpret_cast<uint32_t *>(L.dst);
    for (uint64_t k = threadIdx.x; k < L.words; k += blockDim.x) dst[k] = src[k];
  }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_patch(const Patch *__restrict__ pt, uint64_t n, uint64_t *__restrict__ vmap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vmap[pt[i].idx] = pt[i].val;
}

uint32_t fill_grid(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 4096)); }

}  // namespace

// ------------------------------------------------------------------- spill
void spill_tables(gem_ctx *c, QueryDev &Qd, uint64_t slot_bytes) {
  hipStream_t s = c->stream;
  const uint64_t slot_entries = slot_bytes / 4;
  uint64_t total = 0;
  for (const TableDev &T : Qd.tab)
    if (T.present) total += T.size * 4;
  // mapped: the cache's load kernel reads blocks straight from it
  Qd.host_vals.alloc(std::max<uint64_t>(total, 4), hipHostMallocMapped);
  unsigned char *dst = Qd.host_vals.as<unsigned char>();
  for (TableDev &T : Qd.tab) {
    if (!T.present) continue;
    const uint64_t nkeys = Qd.ncan[T.key_u];
    std::vector<uint64_t> off(nkeys + 1);
    GEM_HIP_CHECK(hipMemcpyAsync(off.data(), T.off.p, 8 * (nkeys + 1), hipMemcpyDeviceToHost, s));
    if (T.size) GEM_HIP_CHECK(hipMemcpyAsync(dst, T.val.p, T.size * 4, hipMemcpyDeviceToHost, s));
    GEM_HIP_CHECK(hipStreamSynchronize(s));
    T.hval = reinterpret_cast<const uint32_t *>(dst);
    dst += T.size * 4;
    c->sstats.table_bytes += T.size * 4;
    uint64_t longest = 0;
    for (uint64_t k = 0; k < nkeys; ++k) longest = std::max(longest, off[k + 1] - off[k]);
    if (longest > slot_entries)
      throw Error{GEM_E_LIMIT, "a table list of " + std::to_string(longest) + " entries is longer than a " +
                                   std::to_string(slot_bytes) + "-byte cache slot"};
    // the widest key range whose every block fits a slot (halving a block
    // that fits gives blocks that fit, and single keys fit)
    uint32_t ks = 20;
    for (;; --ks) {
      bool ok = true;
      for (uint64_t k0 = 0; k0 < nkeys && ok; k0 += 1ull << ks)
        ok = off[std::min<uint64_t>(k0 + (1ull << ks), nkeys)] - off[k0] <= slot_entries;
      if (ok || ks == 0) break;
    }
    T.kshift = ks;
    const uint64_t nblk = (nkeys + (1ull << ks) - 1) >> ks;
    T.blk_off.resize(nblk + 1);
    for (uint64_t b = 0; b <= nblk; ++b) T.blk_off[b] = off[std::min<uint64_t>(b << ks, nkeys)];
    T.val.reset();
  }
  Qd.streamed = true;
}

// ------------------------------------------------------------------- cache
TableCache::TableCache(gem_ctx *c, uint64_t cache_bytes, uint64_t slot_bytes)
    : c_(c), cache_bytes_(cache_bytes) {
  sbytes_[0] = std::min<uint64_t>(SMALL_SLOT, slot_bytes);
  sbytes_[1] = slot_bytes;
  GEM_HIP_CHECK(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
  GEM_HIP_CHECK(hipEventCreate(&ev_step0_));
}

TableCache::~TableCache() {
  if (copy_) {
    (void)hipStreamSynchronize(copy_);
    (void)hipStreamDestroy(copy_);
  }
  if (c_->stream) (void)hipStreamSynchronize(c_->stream);
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  if (ev_step0_) (void)hipEventDestroy(ev_step0_);
}

uint64_t TableCache::guard_addr() const {
  return (uint64_t)(uintptr_t)(pool_.as<unsigned char>() + pool_bytes_ - sbytes_[1]);
}
uint64_t TableCache::slot_addr(int cls, uint32_t s) const {
  return (uint64_t)(uintptr_t)(pool_.as<unsigned char>() + cbase_[cls] + (uint64_t)s * sbytes_[cls]);
}

void TableCache::attach(std::vector<VQueryState *> &qs) {
  hipStream_t s = c_->stream;
  // global block ids, table by table; blocks keep the device-visible address
  // of their values in pinned host memory
  std::vector<std::vector<uint64_t>> &base = base_;
  base.assign(qs.size(), {});
  nblk_ = 0;
  for (size_t i = 0; i < qs.size(); ++i) {
    const QueryDev &Qd = *qs[i]->qd;
    if (!Qd.streamed) throw Error{GEM_E_STATE, "streamed join over tables that were not spilled"};
    void *dev_vals = nullptr;
    GEM_HIP_CHECK(hipHostGetDevicePointer(&dev_vals, Qd.host_vals.p, 0));
    base[i].assign(Qd.tab.size(), 0);
    for (size_t d = 0; d < Qd.tab.size(); ++d) {
      const TableDev &T = Qd.tab[d];
      if (!T.present) continue;
      base[i][d] = nblk_;
      const uint64_t nb = T.blk_off.size() - 1;
      const uint64_t toff = (uint64_t)(reinterpret_cast<const unsigned char *>(T.hval) -
                                       Qd.host_vals.as<const unsigned char>());
      for (uint64_t b = 0; b < nb; ++b) {
        bsrc_.push_back((uint64_t)(uintptr_t)dev_vals + toff + 4 * T.blk_off[b]);
        boff_.push_back(T.blk_off[b]);
        blen_.push_back(T.blk_off[b + 1] - T.blk_off[b]);
      }
      nblk_ += nb;
    }
  }
  if (nblk_ >= NONE32) throw Error{GEM_E_LIMIT, "too many table blocks for one batch"};
  // size classes: blocks of at most SMALL_SLOT bytes take small slots; the
  // cache is split between the classes in proportion to their table bytes,
  // after each class has two slots of its own (with slots of at most
  // SMALL_SLOT bytes every block is in the small class)
  cls_.assign(nblk_, 0);
  uint64_t cbytes[2] = {0, 0}, cblocks[2] = {0, 0};
  for (uint64_t b = 0; b < nblk_; ++b) {
    cls_[b] = 4 * blen_[b] <= sbytes_[0] ? 0 : 1;
    cbytes[cls_[b]] += 4 * blen_[b];
    cblocks[cls_[b]] += 1;
  }
  uint64_t need_slots[2], rest = cache_bytes_;
  for (int k = 0; k < 2; ++k) {
    need_slots[k] = std::min<uint64_t>(cblocks[k], 2);
    if (need_slots[k] * sbytes_[k] > rest) throw Error{GEM_E_LIMIT, "the device table cache is too small"};
    rest -= need_slots[k] * sbytes_[k];
  }
  const double f0 = cbytes[0] + cbytes[1] ? (double)cbytes[0] / (double)(cbytes[0] + cbytes[1]) : 0.5;
  for (int k = 0; k < 2; ++k) {
    const double share = k == 0 ? f0 : 1.0 - f0;
    nslots_[k] = (uint32_t)std::min<uint64_t>(need_slots[k] + (uint64_t)(share * (double)rest) / sbytes_[k],
                                               std::max<uint64_t>(cblocks[k], 1));
  }
  cbase_[0] = 0;
  cbase_[1] = (uint64_t)nslots_[0] * sbytes_[0];
  pool_bytes_ = cbase_[1] + (uint64_t)nslots_[1] * sbytes_[1] + sbytes_[1];  // + the guard slot
  pool_.alloc(pool_bytes_, s);
  GEM_HIP_CHECK(hipMemsetAsync(pool_.as<unsigned char>() + pool_bytes_ - sbytes_[1], 0, sbytes_[1], s));
  for (int k = 0; k < 2; ++k) {
    slot_blk_[k].assign(nslots_[k], NONE64);
    stamp_[k].assign(nslots_[k], 0);
  }
  blk_slot_.assign(nblk_, NONE32);
  live_in_.assign(nblk_, -1);
  const uint64_t n = std::max<uint64_t>(nblk_, 1);
  vmap_.alloc(8 * n, s);
  first_.alloc(4 * n, s);
  last_.alloc(4 * n, s);
  h_first_.alloc(4 * n);
  h_last_.alloc(4 * n);
  std::vector<uint64_t> hv(n, 0);
  for (uint64_t b = 0; b < nblk_; ++b) hv[b] = guard_addr() - 4 * boff_[b];
  GEM_HIP_CHECK(hipMemcpyAsync(vmap_.p, hv.data(), 8 * n, hipMemcpyHostToDevice, s));
  // level descriptors read the map
  for (size_t i = 0; i < qs.size(); ++i) bind(*qs[i], i);
  // the copy stream uses buffers allocated on the context stream
  GEM_HIP_CHECK(hipStreamSynchronize(s));
}

void TableCache::bind(VQueryState &S, size_t i) {
  const QueryDev &Qd = *S.qd;
  if (i >= base_.size()) throw Error{GEM_E_STATE, "table cache: query not attached"};
  for (uint32_t lv = 1; lv < Qd.Q.nv; ++lv) {
    const gemh::VLevelPlan &lp = S.plan.levels[lv];
    VLevel &L = S.hlv[lv];
    for (uint32_t j = 0; j < lp.nb; ++j) {
      const uint32_t d = lp.directed[j];
      L.vmap[j] = vmap_.as<uint64_t>() + base_[i][d];
      L.kshift[j] = Qd.tab[d].kshift;
      L.gblk[j] = (uint32_t)base_[i][d];
    }
  }
  GEM_HIP_CHECK(hipMemcpyAsync(S.levels.p, S.hlv.data(), sizeof(VLevel) * S.hlv.size(), hipMemcpyHostToDevice,
                               c_->stream));
}

void TableCache::reset_marks() {
  if (!nblk_) return;
  hipStream_t s = c_->stream;
  k_fill_u32<<<fill_grid(nblk_), 256, 0, s>>>(first_.as<uint32_t>(), nblk_, NONE32);
  k_fill_u32<<<fill_grid(nblk_), 256, 0, s>>>(last_.as<uint32_t>(), nblk_, 0);
  GEM_LAUNCH_CHECK("k_fill_u32");
}

bool TableCache::plan(uint64_t total) {
  hipStream_t s = c_->stream;
  if (nblk_) {
    GEM_HIP_CHECK(hipMemcpyAsync(h_first_.p, first_.p, 4 * nblk_, hipMemcpyDeviceToHost, s));
    GEM_HIP_CHECK(hipMemcpyAsync(h_last_.p, last_.p, 4 * nblk_, hipMemcpyDeviceToHost, s));
  }
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  collect();  // the previous window's groups have finished
  hits_ = 0;
  if (total == 0) return true;
  const uint32_t *fi = h_first_.as<uint32_t>(), *la = h_last_.as<uint32_t>();
  // needed blocks (empty blocks are never copied: their lists read nothing),
  // counted per size class
  std::vector<uint64_t> need;
  std::vector<uint32_t> scount[2], ecount[2];
  for (int k = 0; k < 2; ++k) {
    scount[k].assign(total + 1, 0);
    ecount[k].assign(total + 1, 0);
  }
  for (uint64_t b = 0; b < nblk_; ++b) {
    if (fi[b] == NONE32 || blen_[b] == 0) continue;
    if (fi[b] >= total || la[b] >= total || la[b] < fi[b]) throw Error{GEM_E_STATE, "block marks out of range"};
    need.push_back(b);
    scount[cls_[b]][fi[b] + 1]++;
    ecount[cls_[b]][la[b] + 1]++;
  }
  // S[t] = #needed blocks first read before task t, E[t] = #last read before t
  for (int k = 0; k < 2; ++k)
    for (uint64_t t = 1; t <= total; ++t) {
      scount[k][t] += scount[k][t - 1];
      ecount[k][t] += ecount[k][t - 1];
    }
  // blocks first read before b and last read at or after lo fit every class
  auto fits = [&](uint64_t b, uint64_t lo) {
    for (int k = 0; k < 2; ++k)
      if ((uint64_t)scount[k][b] - ecount[k][lo] > nslots_[k]) return false;
    return true;
  };
  // groups: the blocks live in the previous and the current group fit the
  // cache together (copies overlap the previous group's kernels); at least
  // four groups when everything fits, so copies still overlap
  groups_.clear();
  const uint64_t target = std::max<uint64_t>(1, (total + 3) / 4);
  uint64_t a_prev = 0, a = 0;
  while (a < total) {
    Group G;
    G.a = a;
    uint64_t lo = a_prev;
    if (!fits(a + 1, lo)) {
      G.serial = true;
      lo = a;
      // the blocks live at task a (used at or before it and at or after it)
      // cannot all be resident: the caller plans a smaller window
      if (!fits(a + 1, a)) {
        groups_.clear();
        return false;
      }
    }
    uint64_t b = a + 1;
    while (b < total && b - a < target && fits(b + 1, lo)) ++b;
    G.b = b;
    groups_.push_back(std::move(G));
    a_prev = a;
    a = b;
  }
  // simulate the cache over the groups: which blocks each group copies and
  // which it evicts (every change is one map patch, every copy one load)
  std::sort(need.begin(), need.end(), [&](uint64_t x, uint64_t y) { return fi[x] < fi[y]; });
  std::vector<Patch> patches;
  std::vector<Load> loads;
  std::vector<uint64_t> live;  // needed blocks first read before the group's end
  size_t nx = 0;
  for (size_t g = 0; g < groups_.size(); ++g) {
    Group &G = groups_[g];
    while (nx < need.size() && fi[need[nx]] < G.b) live.push_back(need[nx++]);
    live.erase(std::remove_if(live.begin(), live.end(), [&](uint64_t b) { return la[b] < G.a; }), live.end());
    const int64_t gid = gbase_ + (int64_t)g;  // unique over the run's iterations
    for (uint64_t b : live) live_in_[b] = gid;
    ++clock_;
    std::vector<uint64_t> miss[2];
    for (uint64_t b : live) {
      if (blk_slot_[b] != NONE32) {
        stamp_[cls_[b]][blk_slot_[b]] = clock_;
        ++hits_;
      } else {
        miss[cls_[b]].push_back(b);
      }
    }
    G.p0 = patches.size();
    G.l0 = loads.size();
    for (int k = 0; k < 2; ++k) {
      if (miss[k].empty()) continue;
      // victims: free slots, then slots whose block is not live in this or
      // (unless serial) the previous group: blocks not read again in this
      // window first (least recently used first), then those read latest
      const int64_t keep_from = G.serial ? gid : gid - 1;
      std::vector<std::pair<uint64_t, uint32_t>> cand;  // (priority, slot)
      for (uint32_t sl = 0; sl < nslots_[k]; ++sl) {
        const uint64_t b = slot_blk_[k][sl];
        if (b == NONE64) {
          cand.push_back({0, sl});
          continue;
        }
        if (live_in_[b] >= keep_from) continue;
        const bool later = fi[b] != NONE32 && la[b] >= G.b;
        cand.push_back({later ? (2ull << 62) - fi[b] : (1ull << 62) + stamp_[k][sl], sl});
      }
      if (cand.size() < miss[k].size()) throw Error{GEM_E_STATE, "table cache plan ran out of slots"};
      std::partial_sort(cand.begin(), cand.begin() + miss[k].size(), cand.end());
      for (size_t m = 0; m < miss[k].size(); ++m) {
        const uint32_t sl = cand[m].second;
        const uint64_t old = slot_blk_[k][sl];
        if (old != NONE64) {
          blk_slot_[old] = NONE32;
          patches.push_back({old, guard_addr() - 4 * boff_[old]});
        }
        const uint64_t b = miss[k][m];
        slot_blk_[k][sl] = b;
        blk_slot_[b] = sl;
        stamp_[k][sl] = clock_;
        patches.push_back({b, slot_addr(k, sl) - 4 * boff_[b]});
        loads.push_back({bsrc_[b], slot_addr(k, sl), blen_[b]});
      }
    }
    G.p1 = patches.size();
    G.l1 = loads.size();
  }
  // one unused id between windows: the previous window's last group is never
  // "the previous group" of this one's first
  gbase_ += (int64_t)groups_.size() + 1;
  // patch and load lists: pinned (async copy source) and device copies;
  // growing them is safe here because the streams are past the last window
  const size_t pbytes = std::max<size_t>(patches.size(), 1) * sizeof(Patch);
  const size_t lbytes = std::max<size_t>(loads.size(), 1) * sizeof(Load);
  if (h_patch_.bytes < pbytes) h_patch_.alloc(pbytes * 2);
  if (h_load_.bytes < lbytes) h_load_.alloc(lbytes * 2);
  bool grown = false;
  if (dpatch_.bytes < pbytes) {
    dpatch_.alloc(pbytes * 2, s);
    grown = true;
  }
  if (dload_.bytes < lbytes) {
    dload_.alloc(lbytes * 2, s);
    grown = true;
  }
  if (grown) GEM_HIP_CHECK(hipStreamSynchronize(s));
  if (!patches.empty()) std::memcpy(h_patch_.p, patches.data(), patches.size() * sizeof(Patch));
  if (!loads.empty()) std::memcpy(h_load_.p, loads.data(), loads.size() * sizeof(Load));
  // events: 3 per group
  while (ev_.size() < 3 * groups_.size()) {
    hipEvent_t e;
    GEM_HIP_CHECK(hipEventCreate(&e));
    ev_.push_back(e);
  }
  GEM_HIP_CHECK(hipEventRecord(ev_step0_, s));
  return true;
}

unsigned int *TableCache::queues() {
  hipStream_t s = c_->stream;
  const size_t need = 4 * 7 * std::max<size_t>(groups_.size(), 1);
  if (queues_.bytes < need) queues_.alloc(need * 2, s);
  GEM_HIP_CHECK(hipMemsetAsync(queues_.p, 0, need, s));
  return queues_.as<unsigned int>();
}

void TableCache::issue(size_t g) {
  const Group &G = groups_[g];
  hipStream_t s = c_->stream;
  // blocks this group may evict were last read by group g-2 (g-1 if serial)
  const int64_t wait_for = G.serial ? (int64_t)g - 1 : (int64_t)g - 2;
  if (g == 0) GEM_HIP_CHECK(hipStreamWaitEvent(copy_, ev_step0_, 0));
  if (wait_for >= 0) GEM_HIP_CHECK(hipStreamWaitEvent(copy_, ev_[3 * wait_for + 2], 0));
  GEM_HIP_CHECK(hipEventRecord(ev_[3 * g], copy_));
  const size_t nl = G.l1 - G.l0;
  if (nl) {
    // the group's blocks, read from pinned host memory by one kernel
    Load *dl = dload_.as<Load>() + G.l0;
    GEM_HIP_CHECK(hipMemcpyAsync(dl, h_load_.as<Load>() + G.l0, nl * sizeof(Load), hipMemcpyHostToDevice, copy_));
    k_load_blocks<<<(uint32_t)std::min<size_t>(nl, 2048), 256, 0, copy_>>>(dl, nl);
    GEM_LAUNCH_CHECK("k_load_blocks");
    for (size_t i = G.l0; i < G.l1; ++i) c_->sstats.h2d_bytes += 4 * h_load_.as<Load>()[i].words;
  }
  c_->sstats.blocks_loaded += nl;
  const size_t np = G.p1 - G.p0;
  if (np) {
    Patch *dp = dpatch_.as<Patch>() + G.p0;
    GEM_HIP_CHECK(hipMemcpyAsync(dp, h_patch_.as<Patch>() + G.p0, np * sizeof(Patch), hipMemcpyHostToDevice, copy_));
    k_patch<<<(uint32_t)((np + 255) / 256), 256, 0, copy_>>>(dp, np, vmap_.as<uint64_t>());
    GEM_LAUNCH_CHECK("k_patch");
  }
  GEM_HIP_CHECK(hipEventRecord(ev_[3 * g + 1], copy_));
  GEM_HIP_CHECK(hipStreamWaitEvent(s, ev_[3 * g + 1], 0));
  c_->sstats.groups += 1;
  c_->sstats.serial_groups += G.serial ? 1 : 0;
}

void TableCache::done(size_t g) { GEM_HIP_CHECK(hipEventRecord(ev_[3 * g + 2], c_->stream)); }

void TableCache::finish() { collect(); }

// Statistics of the planned window's groups once both streams are past them;
// then the plan is dropped.
void TableCache::collect() {
  c_->sstats.blocks_hit += hits_;
  hits_ = 0;
  if (groups_.empty()) return;
  GEM_HIP_CHECK(hipStreamSynchronize(copy_));
  GEM_HIP_CHECK(hipStreamSynchronize(c_->stream));
  for (size_t g = 0; g < groups_.size(); ++g) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_[3 * g], ev_[3 * g + 1]) == hipSuccess) c_->sstats.copy_ms += ms;
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, ev_step0_, ev_[3 * (groups_.size() - 1) + 2]) == hipSuccess) c_->sstats.step_ms += ms;
  groups_.clear();
}

}  // namespace gemd
