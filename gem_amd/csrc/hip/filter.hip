// Candidate filter and position index.
//
// Predicate (out-of-core/lib/filter/src/filter.cu:160-240; in-core
// filter_kernel.cu:198-265): v is a candidate of query vertex u iff
//   lab(v) == lab(u),  deg(u) <= deg(v) + T,
//   sum over query labels l of max(0, nlc_u[l] - nlc_v[l]) <= T,
// where nlc_x[l] counts neighbours of x with label l, over labels present in
// the query.
//
// Layout: one wave owns a tile of 64 consecutive data vertices (lane = vertex)
// and writes, for every query vertex u, one 64-bit word of u's candidate bitmap
// (ballot of the lanes that pass) plus that word's popcount.  A per-u scan of
// the popcounts gives rank[u][tile] = #candidates before the tile, so
//   pos_u(v) = rank[u][v/64] + popcount(bitmap[u][v/64] & below(v % 64))
// equals lower_bound(Can(u), v): the value the reference's cuckoo hash returns
// (unifiedTrie.cuh:84-125, unifiedTrie.cu:54-101), in one cache line, with no
// build retries.  Can(u) is then compacted from the bitmap in ascending order.
//
// Neighbour-label counts: a vertex of degree <= 64 is counted by its own lane
// into a private LDS column (label-major, lane-minor: conflict-free); a
// higher-degree vertex is counted by the whole wave, 64 neighbours per step,
// one ballot + popcount per query label (no atomics).
#include <mutex>
#include <string>

#include "ctx.h"
#include "wave.h"

namespace gemd {

constexpr uint32_t FILTER_WAVES = 4;
constexpr uint32_t LIGHT_DEG = 64;

template <class LabT>
__global__ void __launch_bounds__(FILTER_WAVES * 64)
k_filter(const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr, const LabT *__restrict__ lab, uint32_t n,
         const FilterQuery *__restrict__ fqp, const uint8_t *__restrict__ slot_of, uint64_t *__restrict__ bitmap,
         uint32_t *__restrict__ tile_count, uint32_t *__restrict__ heavy_list, uint32_t *__restrict__ heavy_n) {
  __shared__ FilterQuery fq;
  __shared__ uint32_t hist[FILTER_WAVES][GEM_MAX_QV][64];
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(fqp);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&fq);
    for (uint32_t i = threadIdx.x; i < sizeof(FilterQuery) / 4; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t tile = blockIdx.x * FILTER_WAVES + w;
  if (tile >= fq.W) return;  // whole wave; no block barrier follows
  const uint64_t v = (uint64_t)tile * 64 + l;
  const uint32_t nslots = fq.nslots, T = fq.T;
  uint32_t (*h)[64] = hist[w];

  uint32_t slot = NO_SLOT;
  uint64_t o0 = 0, deg = 0;
  uint32_t cmask = 0;
  if (v < n) {
    slot = slot_of[lab[v]];
    if (slot != NO_SLOT) {
      o0 = off[v];
      deg = off[v + 1] - o0;
      uint32_t g = fq.group[slot];
      while (g) {
        uint32_t u = __builtin_ctz(g);
        g &= g - 1;
        if ((uint64_t)fq.qdeg[u] <= deg + T) cmask |= 1u << u;
      }
    }
  }
  for (uint32_t s = 0; s < nslots; ++s) h[s][l] = 0;
  uint32_t pass = 0;
  // light vertices: own lane, private LDS column
  if (cmask && deg <= LIGHT_DEG) {
    // four neighbours at a time: four independent neighbour -> label -> slot
    // chains in flight instead of one
    uint64_t j = 0;
    for (; j + 4 <= deg; j += 4) {
      const uint32_t a0 = nbr[o0 + j], a1 = nbr[o0 + j + 1], a2 = nbr[o0 + j + 2], a3 = nbr[o0 + j + 3];
      const uint32_t b0 = lab[a0], b1 = lab[a1], b2 = lab[a2], b3 = lab[a3];
      const uint32_t s0 = slot_of[b0], s1 = slot_of[b1], s2 = slot_of[b2], s3 = slot_of[b3];
      if (s0 != NO_SLOT) h[s0][l] += 1;
      if (s1 != NO_SLOT) h[s1][l] += 1;
      if (s2 != NO_SLOT) h[s2][l] += 1;
      if (s3 != NO_SLOT) h[s3][l] += 1;
    }
    for (; j < deg; ++j) {
      uint32_t s = slot_of[lab[nbr[o0 + j]]];
      if (s != NO_SLOT) h[s][l] += 1;
    }
    uint32_t g = cmask;
    while (g) {
      uint32_t u = __builtin_ctz(g);
      g &= g - 1;
      uint32_t deficit = 0;
      for (uint32_t r = 0; r < fq.nreq[u]; ++r) {
        uint32_t need = fq.req_cnt[u][r], have = h[fq.req_slot[u][r]][l];
        deficit += need > have ? need - have : 0;
      }
      if (deficit <= T) pass |= 1u << u;
    }
  }
  // heavy vertices are deferred to k_filter_heavy (one wave each), so a tile
  // full of hubs does not hold its wave for thousands of steps
  const uint64_t heavy = __ballot(cmask != 0 && deg > LIGHT_DEG);
  if (heavy) {
    uint32_t base = 0;
    if (l == 0) base = atomicAdd(heavy_n, (uint32_t)__popcll(heavy));
    base = readlane_u32(base, 0);
    if ((heavy >> l) & 1ull) heavy_list[base + rank_below(heavy)] = (uint32_t)v;
  }
  const uint32_t W = fq.W;
  for (uint32_t u = 0; u < fq.nv; ++u) {
    uint64_t m = __ballot((pass >> u) & 1u);
    if (l == 0) {
      bitmap[(uint64_t)u * W + tile] = m;
      tile_count[(uint64_t)u * W + tile] = (uint32_t)__popcll(m);
    }
  }
}

// Query vertex u's decision for data vertex v of degree deg whose neighbour
// label counts (per query label slot) are cnt: lane u sets v's bit.
__device__ __forceinline__ void filter_decide(const FilterQuery &fq, uint32_t l, uint32_t v, uint64_t deg, uint32_t g,
                                              const uint32_t *cnt, unsigned long long *bitmap,
                                              uint32_t *tile_count) {
  if (l < fq.nv && ((g >> l) & 1u) && (uint64_t)fq.qdeg[l] <= deg + fq.T) {
    uint32_t deficit = 0;
    for (uint32_t r = 0; r < fq.nreq[l]; ++r) {
      const uint32_t need = fq.req_cnt[l][r], have = cnt[fq.req_slot[l][r]];
      deficit += need > have ? need - have : 0;
    }
    if (deficit <= fq.T) {
      atomicOr(&bitmap[(uint64_t)l * fq.W + (v >> 6)], 1ull << (v & 63));
      atomicAdd(&tile_count[(uint64_t)l * fq.W + (v >> 6)], 1u);
    }
  }
}

// Vertices of degree in (LIGHT_DEG, FILTER_HUB_DEG]: one wave per vertex
// (persistent grid over the list k_filter built); larger ones are passed on to
// the hub list for k_filter_hub.  Neighbour labels are counted 64 at a time,
// one ballot + popcount per query label (lane s keeps label s's count); lane u
// then decides query vertex u and sets its bit with an atomic OR.
template <class LabT>
__global__ void __launch_bounds__(FILTER_WAVES * 64)
k_filter_heavy(const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr, const LabT *__restrict__ lab,
               const FilterQuery *__restrict__ fqp, const uint8_t *__restrict__ slot_of,
               unsigned long long *__restrict__ bitmap, uint32_t *__restrict__ tile_count,
               const uint32_t *__restrict__ heavy_list, const uint32_t *__restrict__ heavy_n,
               uint32_t *__restrict__ hub_list, uint32_t *__restrict__ hub_n, uint32_t hub_cap) {
  __shared__ FilterQuery fq;
  __shared__ uint32_t cnt[FILTER_WAVES][GEM_MAX_QV];
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(fqp);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&fq);
    for (uint32_t i = threadIdx.x; i < sizeof(FilterQuery) / 4; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t nh = *heavy_n;
  const uint32_t nslots = fq.nslots;
  for (uint32_t i = blockIdx.x * FILTER_WAVES + w; i < nh; i += gridDim.x * FILTER_WAVES) {
    const uint32_t v = heavy_list[i];
    const uint64_t o0 = off[v], deg = off[v + 1] - o0;
    if (deg > FILTER_HUB_DEG) {
      if (l == 0) {
        const uint32_t h = atomicAdd(hub_n, 1u);
        if (h < hub_cap) hub_list[h] = v;  // hub_cap = all such vertices of the graph
      }
      continue;
    }
    const uint32_t g = fq.group[slot_of[lab[v]]];
    uint32_t acc = 0;
    for (uint64_t j = 0; j < deg; j += 128) {
      // two 64-entry groups: both neighbour -> label -> slot chains in flight
      const bool i0 = j + l < deg, i1 = j + 64 + l < deg;
      const uint32_t a0 = i0 ? nbr[o0 + j + l] : 0, a1 = i1 ? nbr[o0 + j + 64 + l] : 0;
      const uint32_t b0 = i0 ? (uint32_t)lab[a0] : 0, b1 = i1 ? (uint32_t)lab[a1] : 0;
      const uint32_t s0 = i0 ? slot_of[b0] : NO_SLOT, s1 = i1 ? slot_of[b1] : NO_SLOT;
      for (uint32_t q = 0; q < nslots; ++q) {
        const uint32_t c = (uint32_t)__popcll(__ballot(s0 == q)) + (uint32_t)__popcll(__ballot(s1 == q));
        if (l == q) acc += c;
      }
    }
    if (l < nslots) cnt[w][l] = acc;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    filter_decide(fq, l, v, deg, g, cnt[w], bitmap, tile_count);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

// Hubs (degree > FILTER_HUB_DEG): task t = (hub t / maxc, chunk t % maxc) of
// FILTER_HUB_DEG adjacency entries; label counts are added into
// hub_cnt[hub][slot] (zeroed by the host).  maxc = chunks of the largest degree.
template <class LabT>
__global__ void __launch_bounds__(FILTER_WAVES * 64)
k_filter_hub(const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr, const LabT *__restrict__ lab,
             const FilterQuery *__restrict__ fqp, const uint8_t *__restrict__ slot_of,
             const uint32_t *__restrict__ hub_list, const uint32_t *__restrict__ hub_n, uint32_t hub_cap,
             uint32_t maxc, uint32_t *__restrict__ hub_cnt) {
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t nslots = fqp->nslots;
  const uint32_t nh = *hub_n < hub_cap ? *hub_n : hub_cap;
  const uint64_t ntask = (uint64_t)nh * maxc;
  for (uint64_t t = (uint64_t)blockIdx.x * FILTER_WAVES + w; t < ntask; t += (uint64_t)gridDim.x * FILTER_WAVES) {
    const uint32_t h = (uint32_t)(t / maxc), ch = (uint32_t)(t % maxc);
    const uint32_t v = hub_list[h];
    const uint64_t o0 = off[v], deg = off[v + 1] - o0;
    const uint64_t j0 = (uint64_t)ch * FILTER_HUB_DEG;
    if (j0 >= deg) continue;
    const uint64_t j1 = j0 + FILTER_HUB_DEG < deg ? j0 + FILTER_HUB_DEG : deg;
    uint32_t acc = 0;
    for (uint64_t j = j0; j < j1; j += 128) {
      const bool i0 = j + l < j1, i1 = j + 64 + l < j1;
      const uint32_t a0 = i0 ? nbr[o0 + j + l] : 0, a1 = i1 ? nbr[o0 + j + 64 + l] : 0;
      const uint32_t b0 = i0 ? (uint32_t)lab[a0] : 0, b1 = i1 ? (uint32_t)lab[a1] : 0;
      const uint32_t s0 = i0 ? slot_of[b0] : NO_SLOT, s1 = i1 ? slot_of[b1] : NO_SLOT;
      for (uint32_t q = 0; q < nslots; ++q) {
        const uint32_t c = (uint32_t)__popcll(__ballot(s0 == q)) + (uint32_t)__popcll(__ballot(s1 == q));
        if (l == q) acc += c;
      }
    }
    if (l < nslots && acc) atomicAdd(&hub_cnt[(uint64_t)h * GEM_MAX_QV + l], acc);
  }
}

// One wave per hub: the filter decision from the summed label counts.
template <class LabT>
__global__ void __launch_bounds__(FILTER_WAVES * 64)
k_filter_hub_decide(const uint64_t *__restrict__ off, const LabT *__restrict__ lab,
                    const FilterQuery *__restrict__ fqp, const uint8_t *__restrict__ slot_of,
                    const uint32_t *__restrict__ hub_list, const uint32_t *__restrict__ hub_n, uint32_t hub_cap,
                    const uint32_t *__restrict__ hub_cnt, unsigned long long *__restrict__ bitmap,
                    uint32_t *__restrict__ tile_count) {
  __shared__ FilterQuery fq;
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(fqp);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&fq);
    for (uint32_t i = threadIdx.x; i < sizeof(FilterQuery) / 4; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint32_t nh = *hub_n < hub_cap ? *hub_n : hub_cap;
  for (uint32_t h = blockIdx.x * FILTER_WAVES + w; h < nh; h += gridDim.x * FILTER_WAVES) {
    const uint32_t v = hub_list[h];
    const uint64_t deg = off[v + 1] - off[v];
    filter_decide(fq, l, v, deg, fq.group[slot_of[lab[v]]], hub_cnt + (uint64_t)h * GEM_MAX_QV, bitmap, tile_count);
  }
}

// Compaction: Can(u) ascending from the bitmap and ranks.
__global__ void __launch_bounds__(256) k_compact(const uint64_t *__restrict__ bitmap, const uint32_t *__restrict__ rank,
                                               uint32_t nv, uint32_t W, const uint64_t *__restrict__ can_off,
                                               uint32_t *__restrict__ can) {
  const uint32_t tile = blockIdx.x * 4 + threadIdx.x / 64, l = threadIdx.x % 64;
  if (tile >= W) return;
  for (uint32_t u = 0; u < nv; ++u) {
    const uint64_t m = bitmap[(uint64_t)u * W + tile];
    if (!m) continue;
    if (m >> l & 1ull) {
      const uint32_t r = rank[(uint64_t)u * (W + 1) + tile] + rank_below(m);
      can[can_off[u] + r] = tile * 64 + l;
    }
  }
}

__global__ void k_position_lookup(const uint64_t *__restrict__ bitmap, const uint32_t *__restrict__ rank, uint32_t W,
                                  uint32_t n, uint32_t u, const uint32_t *__restrict__ vs, uint64_t cnt,
                                  uint32_t *__restrict__ pos) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t v = vs[i];
  uint32_t p = NOT_FOUND;
  if (v < n) {
    const uint64_t m = bitmap[(uint64_t)u * W + (v >> 6)];
    const uint32_t b = v & 63;
    if (m >> b & 1ull) p = rank[(uint64_t)u * (W + 1) + (v >> 6)] + (uint32_t)__popcll(m & ((1ull << b) - 1));
  }
  pos[i] = p;
}

// Device buffers of one filter run (library-owned, or carved from the
// no-allocation scratch array).
struct FilterBufs {
  FilterQuery *fq = nullptr;
  uint8_t *slot_of = nullptr;
  uint64_t *bitmap = nullptr;      // u64[nv * W]
  uint32_t *tile_count = nullptr;  // u32[nv * W]
  uint32_t *rank = nullptr;        // u32[nv * (W + 1)]
  uint32_t *can = nullptr;         // candidate lists at can_off[u]
  const uint64_t *can_off = nullptr;
  uint32_t *heavy_list = nullptr;  // u32[n]
  uint32_t *heavy_n = nullptr;     // u32
  uint32_t *hub_list = nullptr;    // u32[hub_cap]
  uint32_t *hub_n = nullptr;       // u32
  uint32_t *hub_cnt = nullptr;     // u32[hub_cap * GEM_MAX_QV]
  uint32_t hub_cap = 1;
};

// Host side of a filter run: the query's label slots and per-vertex label
// requirements, the dense-label -> slot map, the candidate-list offsets (upper
// bound per list = frequency of the vertex's label) and the adjacency slots of
// label-matching vertices (for the algorithmic bytes).
struct FilterPrep {
  FilterQuery fq{};
  std::vector<uint8_t> slot_of;
  std::vector<uint64_t> can_off;
  double adj = 0;
};

static void filter_prepare(const GraphDev &G, const gemh::QueryInfo &Q, uint32_t T, uint32_t W, FilterPrep *P) {
  const uint32_t nv = Q.nv;
  // query label slots in first-seen order (in-core/lib/io/src/io.cu:135-161)
  FilterQuery &fq = P->fq;
  fq = FilterQuery{};
  fq.nv = nv;
  fq.T = T;
  fq.W = W;
  std::vector<uint32_t> slot(nv);
  std::vector<uint32_t> slot_label;
  for (uint32_t u = 0; u < nv; ++u) {
    uint32_t s = 0;
    while (s < slot_label.size() && slot_label[s] != Q.lab[u]) ++s;
    if (s == slot_label.size()) slot_label.push_back(Q.lab[u]);
    slot[u] = s;
    fq.group[s] |= 1u << u;
    fq.qdeg[u] = Q.deg[u];
  }
  fq.nslots = (uint32_t)slot_label.size();
  for (uint32_t u = 0; u < nv; ++u) {
    uint32_t cnt[GEM_MAX_QV] = {};
    for (auto [w, e] : Q.adj[u]) cnt[slot[w]]++;
    for (uint32_t s = 0; s < fq.nslots; ++s)
      if (cnt[s]) {
        fq.req_slot[u][fq.nreq[u]] = (uint8_t)s;
        fq.req_cnt[u][fq.nreq[u]] = (uint8_t)cnt[s];
        fq.nreq[u]++;
      }
  }
  // dense data label -> query slot
  const uint32_t nlab = (uint32_t)G.label_freq.size();
  P->slot_of.assign(std::max<uint32_t>(nlab, 1), (uint8_t)NO_SLOT);
  for (uint32_t s = 0; s < fq.nslots; ++s) {
    auto it = G.label_id.find(slot_label[s]);
    if (it != G.label_id.end()) P->slot_of[it->second] = (uint8_t)s;
  }
  // candidate storage: upper bound = frequency of u's label
  P->can_off.assign(nv + 1, 0);
  for (uint32_t u = 0; u < nv; ++u) {
    auto it = G.label_id.find(Q.lab[u]);
    uint64_t ub = it == G.label_id.end() ? 0 : G.label_freq[it->second];
    P->can_off[u + 1] = P->can_off[u] + ub;
  }
  P->adj = 0;
  for (uint32_t sl = 0; sl < fq.nslots; ++sl) {
    auto it = G.label_id.find(slot_label[sl]);
    if (it != G.label_id.end()) P->adj += (double)G.label_deg[it->second];
  }
}

// |Can(u)| = rank[u][W] for every u.
__global__ void k_gather_counts(const uint32_t *__restrict__ rank, uint32_t nv, uint32_t W, uint32_t *__restrict__ out) {
  for (uint32_t u = threadIdx.x; u < nv; u += blockDim.x) out[u] = rank[(uint64_t)u * (W + 1) + W];
}

__global__ void k_zero_u32(uint32_t *p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// The filter's kernels on the context stream: per-tile bitmaps and counts
// (light vertices in k_filter, heavy ones in k_filter_heavy, hubs in chunks),
// the rank scan, and the compaction of Can(u).  No allocation; the counters
// are cleared by a kernel (no memset node when a caller captures the stream).
static void filter_launch(gem_ctx *c, uint32_t nv, uint32_t W, const FilterBufs &B) {
  const GraphDev &G = graph(c);
  hipStream_t s = c->stream;
  k_zero_u32<<<1, 64, 0, s>>>(B.heavy_n, 1);
  k_zero_u32<<<1, 64, 0, s>>>(B.hub_n, 1);
  if (G.n_hub)
    k_zero_u32<<<(uint32_t)std::min<uint64_t>(((uint64_t)B.hub_cap * GEM_MAX_QV + 255) / 256, 1024), 256, 0, s>>>(
        B.hub_cnt, (uint64_t)B.hub_cap * GEM_MAX_QV);
  GEM_LAUNCH_CHECK("k_zero_u32");
  const uint32_t blocks = (W + FILTER_WAVES - 1) / FILTER_WAVES;
  {
    ProfLaunch pl(c, K_FILTER);
    if (G.lab8)
      k_filter<uint8_t><<<blocks, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.nbr_dev(),
                                                             G.lab.as<uint8_t>(), G.n, B.fq, B.slot_of, B.bitmap,
                                                             B.tile_count, B.heavy_list, B.heavy_n);
    else
      k_filter<uint32_t><<<blocks, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.nbr_dev(),
                                                              G.lab.as<uint32_t>(), G.n, B.fq, B.slot_of, B.bitmap,
                                                              B.tile_count, B.heavy_list, B.heavy_n);
    GEM_LAUNCH_CHECK("k_filter");
  }
  const uint32_t maxc = (uint32_t)((G.max_deg + FILTER_HUB_DEG - 1) / FILTER_HUB_DEG);
  auto *bm = reinterpret_cast<unsigned long long *>(B.bitmap);
  {
    ProfLaunch pl(c, K_FILTER_HEAVY);
    const uint32_t hb = 1024;  // persistent: 4096 waves walk the heavy list
    if (G.lab8)
      k_filter_heavy<uint8_t><<<hb, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.nbr_dev(),
                                                               G.lab.as<uint8_t>(), B.fq, B.slot_of, bm, B.tile_count,
                                                               B.heavy_list, B.heavy_n, B.hub_list, B.hub_n,
                                                               B.hub_cap);
    else
      k_filter_heavy<uint32_t><<<hb, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.nbr_dev(),
                                                                G.lab.as<uint32_t>(), B.fq, B.slot_of, bm,
                                                                B.tile_count, B.heavy_list, B.heavy_n, B.hub_list,
                                                                B.hub_n, B.hub_cap);
    GEM_LAUNCH_CHECK("k_filter_heavy");
    if (G.n_hub) {
      const uint32_t tb = (uint32_t)std::min<uint64_t>(((uint64_t)G.n_hub * maxc + FILTER_WAVES - 1) / FILTER_WAVES,
                                                       2048);
      const uint32_t db = (G.n_hub + FILTER_WAVES - 1) / FILTER_WAVES;
      if (G.lab8) {
        k_filter_hub<uint8_t><<<tb, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.nbr_dev(),
                                                               G.lab.as<uint8_t>(), B.fq, B.slot_of, B.hub_list,
                                                               B.hub_n, B.hub_cap, maxc, B.hub_cnt);
        k_filter_hub_decide<uint8_t><<<db, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.lab.as<uint8_t>(), B.fq,
                                                                      B.slot_of, B.hub_list, B.hub_n, B.hub_cap,
                                                                      B.hub_cnt, bm, B.tile_count);
      } else {
        k_filter_hub<uint32_t><<<tb, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.nbr_dev(),
                                                                G.lab.as<uint32_t>(), B.fq, B.slot_of, B.hub_list,
                                                                B.hub_n, B.hub_cap, maxc, B.hub_cnt);
        k_filter_hub_decide<uint32_t><<<db, FILTER_WAVES * 64, 0, s>>>(G.off.as<uint64_t>(), G.lab.as<uint32_t>(),
                                                                       B.fq, B.slot_of, B.hub_list, B.hub_n,
                                                                       B.hub_cap, B.hub_cnt, bm, B.tile_count);
      }
      GEM_LAUNCH_CHECK("k_filter_hub");
    }
  }
  scan_segments_u32(c, B.tile_count, B.rank, nv, W);
  if (B.can) {
    ProfLaunch pl(c, K_COMPACT);
    k_compact<<<(W + 3) / 4, 256, 0, s>>>(B.bitmap, B.rank, nv, W, B.can_off, B.can);
    GEM_LAUNCH_CHECK("k_compact");
  }
}

void run_filter(gem_ctx *c, const gem_query_view *qv, uint32_t T) { run_filter(c, c->q, qv, T); }

void run_filter(gem_ctx *c, QueryDev &Qd, const gem_query_view *qv, uint32_t T) {
  if (!graph(c).loaded) throw Error{GEM_E_STATE, "no data graph loaded"};
  Qd = QueryDev();
  std::string why;
  int st = gemh::analyse_query(qv, &Qd.Q, &why);
  if (st) throw Error{st, why};
  const gemh::QueryInfo &Q = Qd.Q;
  const GraphDev &G = graph(c);
  const uint32_t nv = Q.nv;
  const uint32_t W = (G.n + 63) / 64;
  Qd.W = W;
  FilterPrep P;
  filter_prepare(G, Q, T, W, &P);
  Qd.can_off = P.can_off;

  hipStream_t s = c->stream;
  GEM_HIP_CHECK(hipEventRecord(c->ev[0], s));
  Qd.fq.alloc(sizeof(FilterQuery), s);
  Qd.slot_of.alloc(P.slot_of.size(), s);
  GEM_HIP_CHECK(hipMemcpyAsync(Qd.fq.p, &P.fq, sizeof(FilterQuery), hipMemcpyHostToDevice, s));
  GEM_HIP_CHECK(hipMemcpyAsync(Qd.slot_of.p, P.slot_of.data(), P.slot_of.size(), hipMemcpyHostToDevice, s));
  Qd.bitmap.alloc((uint64_t)nv * W * 8, s);
  DevBuf tile_count;
  tile_count.alloc((uint64_t)nv * W * 4, s);
  Qd.rank.alloc((uint64_t)nv * (W + 1) * 4, s);
  Qd.can.alloc(std::max<uint64_t>(Qd.can_off[nv], 1) * 4, s);
  DevBuf can_off;
  can_off.alloc((nv + 1) * 8, s);
  GEM_HIP_CHECK(hipMemcpyAsync(can_off.p, Qd.can_off.data(), (nv + 1) * 8, hipMemcpyHostToDevice, s));
  DevBuf heavy_list, heavy_n, hub_list, hub_n, hub_cnt;
  heavy_list.alloc(std::max<uint64_t>(G.n, 1) * 4, s);
  heavy_n.alloc(4, s);
  const uint32_t hub_cap = std::max<uint32_t>(G.n_hub, 1);
  hub_list.alloc((uint64_t)hub_cap * 4, s);
  hub_n.alloc(4, s);
  hub_cnt.alloc((uint64_t)hub_cap * GEM_MAX_QV * 4, s);
  FilterBufs B;
  B.fq = Qd.fq.as<FilterQuery>();
  B.slot_of = Qd.slot_of.as<uint8_t>();
  B.bitmap = Qd.bitmap.as<uint64_t>();
  B.tile_count = tile_count.as<uint32_t>();
  B.rank = Qd.rank.as<uint32_t>();
  B.can = Qd.can.as<uint32_t>();
  B.can_off = can_off.as<uint64_t>();
  B.heavy_list = heavy_list.as<uint32_t>();
  B.heavy_n = heavy_n.as<uint32_t>();
  B.hub_list = hub_list.as<uint32_t>();
  B.hub_n = hub_n.as<uint32_t>();
  B.hub_cnt = hub_cnt.as<uint32_t>();
  B.hub_cap = hub_cap;
  filter_launch(c, nv, W, B);
  // algorithmic bytes: 16 B per vertex (label, degree, offset pair), 8 B per
  // adjacency slot of a label-matching vertex (neighbour id + its label), the
  // candidate bitmaps and per-tile counts written
  prof_bytes(c, K_FILTER, 16.0 * G.n + 8.0 * P.adj + nv * (double)W * 8 + nv * (double)W * 4);
  // |Can(u)| = rank[u][W], gathered by one kernel and copied back at once
  DevBuf dn;
  dn.alloc(4ull * nv, s);
  k_gather_counts<<<1, 64, 0, s>>>(Qd.rank.as<uint32_t>(), nv, W, dn.as<uint32_t>());
  GEM_LAUNCH_CHECK("k_gather_counts");
  uint32_t *h = pinned<uint32_t>(c, nv);
  GEM_HIP_CHECK(hipMemcpyAsync(h, dn.p, 4ull * nv, hipMemcpyDeviceToHost, s));
  GEM_HIP_CHECK(hipEventRecord(c->ev[1], s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  Qd.ncan.assign(h, h + nv);
  double ncan_sum = 0;
  for (uint32_t u = 0; u < nv; ++u) ncan_sum += Qd.ncan[u];
  prof_bytes(c, K_COMPACT, nv * (double)W * 12 + 4.0 * ncan_sum);
  prof_flush(c);
  float ms = 0;
  GEM_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->times.filter_ms = ms;
  c->times.index_ms = 0;  // the index (bitmap + rank) is a by-product of the filter
  Qd.filtered = true;
}

// ------------------------------------------------ no-allocation variant
// Scratch of gem_hip_filter_noalloc: the query description, the label map,
// candidate-list offsets, bitmaps, tile counts, ranks and the heavy/hub lists
// of one call, carved from one fixed device array (each device holds its own
// copy until the process ends).  Calls that overlap on one device share it and
// race.
constexpr uint64_t FILTER_SCRATCH_BYTES = 64ull << 20;
constexpr int SCRATCH_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[FILTER_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[SCRATCH_MAX_DEVICES];  // device address of `scratch`, per device

void filter_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) throw Error{GEM_E_LIMIT, "device id beyond the scratch table"};
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error{GEM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
  g_scratch_ptr[device] = static_cast<unsigned char *>(p);
}

uint64_t filter_scratch_bytes(const GraphDev &G, uint32_t nv) {
  const uint64_t W = (G.n + 63) / 64;
  const uint64_t hub_cap = std::max<uint32_t>(G.n_hub, 1);
  auto up = [](uint64_t x) { return (x + 255) / 256 * 256; };
  return up(sizeof(FilterQuery)) + up(std::max<uint64_t>(G.label_freq.size(), 1)) + up(8ull * (nv + 1)) +
         up(8ull * nv * W) + up(4ull * nv * W) + up(4ull * nv * (W + 1)) + up(4ull * std::max<uint64_t>(G.n, 1)) +
         up(4) + up(4 * hub_cap) + up(4) + up(4ull * hub_cap * GEM_MAX_QV);
}

void run_filter_noalloc(gem_ctx *c, const gem_query_view *qv, uint32_t T, uint32_t *d_num_can, uint32_t *d_can,
                        uint64_t can_pitch) {
  if (!graph(c).loaded) throw Error{GEM_E_STATE, "no data graph loaded"};
  if (!d_num_can) throw Error{GEM_E_INVALID, "d_num_can is null"};
  if (c->device < 0 || c->device >= SCRATCH_MAX_DEVICES || !g_scratch_ptr[c->device])
    throw Error{GEM_E_STATE, "scratch not initialised on this device"};
  gemh::QueryInfo Q;
  std::string why;
  int st = gemh::analyse_query(qv, &Q, &why);
  if (st) throw Error{st, why};
  const GraphDev &G = graph(c);
  const uint32_t nv = Q.nv;
  const uint32_t W = (G.n + 63) / 64;
  const uint64_t need = filter_scratch_bytes(G, nv);
  if (need > FILTER_SCRATCH_BYTES)
    throw Error{GEM_E_LIMIT, "gem_hip_filter_noalloc: this graph and query need " + std::to_string(need) +
                                 " bytes of scratch, more than its " + std::to_string(FILTER_SCRATCH_BYTES)};
  FilterPrep P;
  filter_prepare(G, Q, T, W, &P);
  // candidate lists at u * can_pitch in the caller's array: each list needs
  // room for its label's frequency (the bound the kernels write within)
  std::vector<uint64_t> can_off(nv + 1, 0);
  for (uint32_t u = 0; u < nv; ++u) {
    const uint64_t ub = P.can_off[u + 1] - P.can_off[u];
    if (d_can && ub > can_pitch)
      throw Error{GEM_E_INVALID, "can_pitch is smaller than the frequency of a query vertex's label"};
    can_off[u] = (uint64_t)u * can_pitch;
  }
  can_off[nv] = (uint64_t)nv * can_pitch;
  // carve the scratch array
  unsigned char *base = g_scratch_ptr[c->device];
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    unsigned char *p = base + at;
    at += (bytes + 255) / 256 * 256;
    return p;
  };
  const uint32_t hub_cap = std::max<uint32_t>(G.n_hub, 1);
  FilterBufs B;
  B.fq = reinterpret_cast<FilterQuery *>(take(sizeof(FilterQuery)));
  B.slot_of = take(P.slot_of.size());
  uint64_t *d_can_off = reinterpret_cast<uint64_t *>(take(8ull * (nv + 1)));
  B.can_off = d_can_off;
  B.bitmap = reinterpret_cast<uint64_t *>(take(8ull * nv * W));
  B.tile_count = reinterpret_cast<uint32_t *>(take(4ull * nv * W));
  B.rank = reinterpret_cast<uint32_t *>(take(4ull * nv * (W + 1)));
  B.heavy_list = reinterpret_cast<uint32_t *>(take(4ull * std::max<uint32_t>(G.n, 1)));
  B.heavy_n = reinterpret_cast<uint32_t *>(take(4));
  B.hub_list = reinterpret_cast<uint32_t *>(take(4ull * hub_cap));
  B.hub_n = reinterpret_cast<uint32_t *>(take(4));
  B.hub_cnt = reinterpret_cast<uint32_t *>(take(4ull * hub_cap * GEM_MAX_QV));
  B.hub_cap = hub_cap;
  if (at > FILTER_SCRATCH_BYTES) throw Error{GEM_E_STATE, "scratch layout exceeds its size"};
  B.can = d_can;  // null: counts only (no compaction)
  hipStream_t s = c->stream;
  GEM_HIP_CHECK(hipMemcpyAsync(B.fq, &P.fq, sizeof(FilterQuery), hipMemcpyHostToDevice, s));
  GEM_HIP_CHECK(hipMemcpyAsync(B.slot_of, P.slot_of.data(), P.slot_of.size(), hipMemcpyHostToDevice, s));
  GEM_HIP_CHECK(hipMemcpyAsync(d_can_off, can_off.data(), 8ull * (nv + 1), hipMemcpyHostToDevice, s));
  filter_launch(c, nv, W, B);
  k_gather_counts<<<1, 64, 0, s>>>(B.rank, nv, W, d_num_can);
  GEM_LAUNCH_CHECK("k_gather_counts");
  // the host arrays above are read by the copies: wait before they go
  GEM_HIP_CHECK(hipStreamSynchronize(s));
}

void run_position_lookup(gem_ctx *c, uint32_t u, const uint32_t *h_vertices, uint64_t n, uint32_t *h_pos) {
  QueryDev &Qd = c->q;
  if (!Qd.filtered) throw Error{GEM_E_STATE, "filter has not run"};
  if (u >= Qd.Q.nv) throw Error{GEM_E_INVALID, "query vertex out of range"};
  if (n == 0) return;
  hipStream_t s = c->stream;
  DevBuf dv, dp;
  dv.alloc(n * 4, s);
  dp.alloc(n * 4, s);
  GEM_HIP_CHECK(hipMemcpyAsync(dv.p, h_vertices, n * 4, hipMemcpyHostToDevice, s));
  k_position_lookup<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(Qd.bitmap.as<uint64_t>(), Qd.rank.as<uint32_t>(), Qd.W,
                                                                c->g.n, u, dv.as<uint32_t>(), n, dp.as<uint32_t>());
  GEM_LAUNCH_CHECK("k_position_lookup");
  GEM_HIP_CHECK(hipMemcpyAsync(h_pos, dp.p, n * 4, hipMemcpyDeviceToHost, s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace gemd
