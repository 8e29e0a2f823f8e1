// C ABI entry points of the device library (declared in include/gem_hip.h).
#include <cstring>
#include <new>
#include <thread>

#include "../host/emb_export.h"
#include "ctx.h"

using namespace gemd;

namespace {

template <class F>
int guard(gem_ctx *c, F &&f) {
  if (!c) return GEM_E_INVALID;
  try {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) throw Error{GEM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)};
    f();
    return GEM_OK;
  } catch (const Error &e) {
    c->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    c->err = "host allocation failed";
    return GEM_E_NOMEM;
  } catch (const std::exception &e) {
    c->err = e.what();
    return GEM_E_INVALID;
  }
}

// Filter, tables and GEM's edge order of q (the plain pipeline before the join).
std::vector<uint32_t> plain_prepare(gem_ctx *c, const gem_query_view *q) {
  run_filter(c, q, 0);
  run_tables(c);
  const gemh::QueryInfo &Q = c->q.Q;
  std::vector<uint64_t> sz(Q.ne);
  for (uint32_t e = 0; e < Q.ne; ++e) {
    const TableDev &a = c->q.tab[2 * e], &b = c->q.tab[2 * e + 1];
    sz[e] = a.present ? a.size : b.size;
  }
  std::vector<uint32_t> order(Q.ne);
  const int st = gemh::gem_order(Q, sz.data(), order.data());
  if (st) throw Error{st, "GEM ordering failed"};
  return order;
}

// memcpy of a large host array over up to 16 threads.
void parallel_copy(void *dst, const void *src, size_t bytes) {
  const size_t nt = std::min<size_t>(16, std::max<size_t>(1, bytes >> 26));
  std::vector<std::thread> th;
  const size_t chunk = (bytes + nt - 1) / nt;
  for (size_t t = 0; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(bytes, b + chunk);
    if (b < e)
      th.emplace_back([=]() { std::memcpy((char *)dst + b, (const char *)src + b, e - b); });
  }
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

int gem_hip_create(int device, gem_ctx **out) {
  if (!out) return GEM_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GEM_E_HIP;
  auto *c = new (std::nothrow) gem_ctx();
  if (!c) return GEM_E_NOMEM;
  c->device = device;
  int st = guard(c, [&]() {
    GEM_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) GEM_HIP_CHECK(hipEventCreate(&e));
    filter_scratch_init(device);
  });
  if (st) {
    gem_hip_destroy(c);
    return st;
  }
  *out = c;
  return GEM_OK;
}

void gem_hip_destroy(gem_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->prep) {
    gem_hip_destroy(c->prep);
    c->prep = nullptr;
  }
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (gem_mlp *m : c->mlps) gemd::mlp_detach(m);
  c->mlps.clear();
  c->q = QueryDev();
  c->g = GraphDev();
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  hipStream_t s = c->stream;
  delete c;  // its remaining device buffers go back to the block cache
  if (s) {
    (void)hipStreamSynchronize(s);
    gemd::pool_release(s);
    (void)hipStreamDestroy(s);
  }
}

const char *gem_hip_last_error(const gem_ctx *c) { return c ? c->err.c_str() : "null context"; }

int gem_hip_set_level_budget(gem_ctx *c, uint64_t bytes) {
  if (!c || bytes < (1ull << 20)) return GEM_E_INVALID;
  c->level_budget_bytes = bytes;
  return GEM_OK;
}

int gem_hip_set_persistent_blocks(gem_ctx *c, uint32_t blocks) {
  if (!c || blocks == 0 || blocks > (1u << 20)) return GEM_E_INVALID;
  c->persistent_blocks = blocks;
  return GEM_OK;
}

int gem_hip_set_heavy_split(gem_ctx *c, uint32_t threshold, uint32_t chunk) {
  if (!c || threshold < 32 || threshold > (1u << 24) || chunk < 64 || chunk > (1u << 20)) return GEM_E_INVALID;
  c->heavy_t = threshold;
  c->heavy_chunk = chunk;
  return GEM_OK;
}

int gem_hip_set_max_active(gem_ctx *c, uint32_t queries) {
  if (!c) return GEM_E_INVALID;
  c->max_active = queries;
  return GEM_OK;
}

int gem_hip_set_level_swap(gem_ctx *c, uint64_t host_bytes) {
  if (!c) return GEM_E_INVALID;
  c->swap_bytes = host_bytes;
  return GEM_OK;
}

int gem_hip_swap_stats(const gem_ctx *c, gem_swap_stats *out) {
  if (!c || !out) return GEM_E_INVALID;
  const gemh::SchedStats &st = c->last_sched;
  *out = gem_swap_stats{st.swaps_out, st.swaps_in, st.swap_bytes_out, st.swap_bytes_in, st.peak_host};
  return GEM_OK;
}

int gem_hip_set_racing(gem_ctx *c, int on) {
  if (!c) return GEM_E_INVALID;
  c->racing = on != 0;
  return GEM_OK;
}

int gem_hip_set_prep_overlap(gem_ctx *c, int on) {
  if (!c) return GEM_E_INVALID;
  c->prep_overlap = on != 0;
  return GEM_OK;
}

int gem_hip_set_relaxed_sharing(gem_ctx *c, int on) {
  if (!c) return GEM_E_INVALID;
  c->relaxed_share = on != 0;
  return GEM_OK;
}

int gem_hip_set_graph_residency(gem_ctx *c, int host) {
  if (!c || host < 0 || host > 1) return GEM_E_INVALID;
  c->graph_host = host;
  return GEM_OK;
}

int gem_hip_set_graph(gem_ctx *c, int on) {
  if (!c) return GEM_E_INVALID;
  c->use_graph = on != 0;
  return GEM_OK;
}

int gem_hip_set_table_stream(gem_ctx *c, uint64_t cache_bytes, uint64_t slot_bytes) {
  if (!c) return GEM_E_INVALID;
  if (slot_bytes == 0) slot_bytes = 1ull << 20;
  // slots hold whole u32 lists and are 256-byte aligned; two slots at least
  if (slot_bytes % 256 || (cache_bytes && cache_bytes < 2 * slot_bytes)) {
    c->err = "slot_bytes must be a multiple of 256 and cache_bytes at least two slots";
    return GEM_E_INVALID;
  }
  c->stream_cache_bytes = cache_bytes;
  c->stream_slot_bytes = slot_bytes;
  return GEM_OK;
}

int gem_hip_stream_stats(gem_ctx *c, gem_stream_stats *out) {
  if (!c || !out) return GEM_E_INVALID;
  *out = c->sstats;
  c->sstats = gem_stream_stats{};
  return GEM_OK;
}

int gem_hip_set_work_limit(gem_ctx *c, uint64_t rows) {
  if (!c) return GEM_E_INVALID;
  c->work_limit_rows = rows;
  return GEM_OK;
}

int gem_hip_load_data_csr(gem_ctx *c, const gem_csr_view *g) {
  return guard(c, [&]() {
    if (!g || !g->offsets || !g->labels || (g->num_slots && !g->neighbors))
      throw Error{GEM_E_INVALID, "null graph array"};
    if (g->offsets[0] != 0 || g->offsets[g->num_v] != g->num_slots)
      throw Error{GEM_E_INVALID, "offsets do not span num_slots"};
    GraphDev &G = c->g;
    c->q = QueryDev();
    G = GraphDev();
    G.n = g->num_v;
    G.slots = g->num_slots;
    for (uint32_t v = 0; v < g->num_v; ++v) {
      if (g->offsets[v + 1] < g->offsets[v]) throw Error{GEM_E_INVALID, "offsets not monotone"};
      G.max_deg = std::max<uint64_t>(G.max_deg, g->offsets[v + 1] - g->offsets[v]);
      if (g->offsets[v + 1] - g->offsets[v] > gemd::FILTER_HUB_DEG) ++G.n_hub;
    }
    // dense label ids in first-seen order
    std::vector<uint32_t> dense(g->num_v);
    for (uint32_t v = 0; v < g->num_v; ++v) {
      auto it = G.label_id.find(g->labels[v]);
      uint32_t id;
      if (it == G.label_id.end()) {
        id = (uint32_t)G.label_id.size();
        G.label_id.emplace(g->labels[v], id);
        G.label_freq.push_back(0);
        G.label_deg.push_back(0);
      } else {
        id = it->second;
      }
      G.label_freq[id]++;
      G.label_deg[id] += g->offsets[v + 1] - g->offsets[v];
      dense[v] = id;
    }
    G.lab8 = G.label_freq.size() <= 255;
    hipStream_t s = c->stream;
    G.off.alloc(((uint64_t)G.n + 1) * 8, s);
    GEM_HIP_CHECK(hipMemcpyAsync(G.off.p, g->offsets, ((uint64_t)G.n + 1) * 8, hipMemcpyHostToDevice, s));
    if (c->graph_host) {
      // Host-resident graph: the neighbour array (all but 12 B per vertex of
      // the CSR) stays in host memory, pinned and mapped; the filter and table
      // kernels read it over PCIe, non-coherent so that L2 may cache it.
      G.nbr_host.alloc(std::max<uint64_t>(G.slots, 1) * 4, hipHostMallocMapped | hipHostMallocNonCoherent);
      parallel_copy(G.nbr_host.p, g->neighbors, G.slots * 4);
      void *dp = nullptr;
      GEM_HIP_CHECK(hipHostGetDevicePointer(&dp, G.nbr_host.p, 0));
      G.nbr_mapped = static_cast<const uint32_t *>(dp);
    } else {
      G.nbr.alloc(std::max<uint64_t>(G.slots, 1) * 4, s);
      if (G.slots) GEM_HIP_CHECK(hipMemcpyAsync(G.nbr.p, g->neighbors, G.slots * 4, hipMemcpyHostToDevice, s));
    }
    if (G.lab8) {
      std::vector<uint8_t> l8(dense.begin(), dense.end());
      G.lab.alloc(std::max<uint64_t>(G.n, 1), s);
      GEM_HIP_CHECK(hipMemcpyAsync(G.lab.p, l8.data(), G.n, hipMemcpyHostToDevice, s));
      GEM_HIP_CHECK(hipStreamSynchronize(s));
    } else {
      G.lab.alloc(std::max<uint64_t>(G.n, 1) * 4, s);
      GEM_HIP_CHECK(hipMemcpyAsync(G.lab.p, dense.data(), (uint64_t)G.n * 4, hipMemcpyHostToDevice, s));
      GEM_HIP_CHECK(hipStreamSynchronize(s));
    }
    G.loaded = true;
  });
}

int gem_hip_filter(gem_ctx *c, const gem_query_view *q, uint32_t threshold, uint32_t *h_num_can, uint32_t *h_can,
                   uint64_t can_pitch) {
  return guard(c, [&]() {
    if (!h_num_can) throw Error{GEM_E_INVALID, "h_num_can is null"};
    run_filter(c, q, threshold);
    const QueryDev &Qd = c->q;
    for (uint32_t u = 0; u < Qd.Q.nv; ++u) h_num_can[u] = Qd.ncan[u];
    if (h_can) {
      for (uint32_t u = 0; u < Qd.Q.nv; ++u) {
        if (Qd.ncan[u] > can_pitch) throw Error{GEM_E_INVALID, "can_pitch smaller than a candidate list"};
        if (Qd.ncan[u])
          GEM_HIP_CHECK(hipMemcpyAsync(h_can + (uint64_t)u * can_pitch, Qd.can.as<uint32_t>() + Qd.can_off[u],
                                       (uint64_t)Qd.ncan[u] * 4, hipMemcpyDeviceToHost, c->stream));
      }
      GEM_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
  });
}

int gem_hip_filter_noalloc(gem_ctx *c, const gem_query_view *q, uint32_t threshold, uint32_t *d_num_can,
                           uint32_t *d_can, uint64_t can_pitch) {
  return guard(c, [&]() {
    run_filter_noalloc(c, q, threshold, d_num_can, d_can, can_pitch);
  });
}

int gem_hip_position_lookup(gem_ctx *c, uint32_t u, const uint32_t *vertices, uint64_t n, uint32_t *pos_out) {
  return guard(c, [&]() {
    if (n && (!vertices || !pos_out)) throw Error{GEM_E_INVALID, "null array"};
    run_position_lookup(c, u, vertices, n, pos_out);
  });
}

int gem_hip_build_edge_tables(gem_ctx *c, uint64_t *h_table_sizes) {
  return guard(c, [&]() {
    run_tables(c);
    if (h_table_sizes)
      for (uint32_t d = 0; d < 2 * c->q.Q.ne; ++d) h_table_sizes[d] = c->q.tab[d].present ? c->q.tab[d].size : 0;
  });
}

int gem_hip_read_edge_table(gem_ctx *c, uint32_t d, uint64_t *offsets, uint32_t *values) {
  return guard(c, [&]() {
    const QueryDev &Qd = c->q;
    if (!Qd.tabled) throw Error{GEM_E_STATE, "edge tables have not been built"};
    if (d >= Qd.tab.size() || !Qd.tab[d].present) throw Error{GEM_E_INVALID, "no table for this directed edge"};
    const TableDev &T = Qd.tab[d];
    const uint64_t nrow = Qd.ncan[T.key_u];
    if (offsets)
      GEM_HIP_CHECK(hipMemcpyAsync(offsets, T.off.p, (nrow + 1) * 8, hipMemcpyDeviceToHost, c->stream));
    if (values && T.size)
      GEM_HIP_CHECK(hipMemcpyAsync(values, T.val.p, T.size * 4, hipMemcpyDeviceToHost, c->stream));
    GEM_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int gem_hip_join(gem_ctx *c, const uint32_t *edge_order, uint64_t cap, gem_join_result *out, gem_step_log *log,
                 uint32_t *emb_out, uint64_t emb_capacity) {
  return guard(c, [&]() {
    if (!edge_order || !out) throw Error{GEM_E_INVALID, "null argument"};
    JoinOut jo;
    run_join(c, edge_order, cap, &jo, log, emb_out, emb_capacity);
    out->count = jo.count;
    out->saturated = jo.saturated ? 1 : 0;
    out->num_steps = c->q.Q.ne - 1;
    out->num_embeddings_written = jo.emb_written;
    out->rows_processed = jo.rows_processed;
    out->work_limited = jo.work_limited ? 1 : 0;
    out->restarts = 0;
    out->work = jo.rows_processed;
  });
}

int gem_hip_memory_stats(gem_ctx *c, int reset_peak, gem_memory_stats *out) {
  return guard(c, [&]() {
    if (!out) throw Error{GEM_E_INVALID, "null argument"};
    const gemh::PoolStats p = pool_stats(reset_peak != 0);
    out->in_use = p.in_use;
    out->peak_in_use = p.peak_in_use;
    out->held = p.held;
    out->reusable = p.empty_segments + p.cached_small;
    out->level_budget = c->last_budget;
    out->peak_charged = c->last_sched.peak_charged;
    out->iterations = c->last_sched.iterations;
    out->forced = c->last_sched.forced;
  });
}

int gem_hip_set_profiling(gem_ctx *c, int on) {
  return guard(c, [&]() {
    GEM_HIP_CHECK(hipStreamSynchronize(c->stream));
    prof_flush(c);
    for (auto &s : c->prof.st) s = KStat();
    c->prof.on = on != 0;
  });
}

int gem_hip_kernel_stats(gem_ctx *c, gem_kernel_stat *out, uint32_t max, uint32_t *n) {
  return guard(c, [&]() {
    if (!out || !n) throw Error{GEM_E_INVALID, "null argument"};
    GEM_HIP_CHECK(hipStreamSynchronize(c->stream));
    prof_flush(c);
    uint32_t k = 0;
    for (int id = 0; id < K_NUM && k < max; ++id) {
      gem_kernel_stat &o = out[k++];
      std::memset(o.name, 0, sizeof(o.name));
      std::strncpy(o.name, kernel_name(id), sizeof(o.name) - 1);
      o.launches = c->prof.st[id].launches;
      o.ms = c->prof.st[id].ms;
      o.bytes = c->prof.st[id].bytes;
    }
    *n = k;
  });
}

int gem_hip_match_batch(gem_ctx *c, const gem_query_view *queries, uint32_t nq, const uint32_t *const *vertex_orders,
                        const gem_batch_opts *opts, gem_join_result *results, uint64_t *level_rows) {
  return guard(c, [&]() {
    if (nq && (!queries || !results)) throw Error{GEM_E_INVALID, "null argument"};
    run_match_batch(c, queries, nq, vertex_orders, opts, results, level_rows);
  });
}

int gem_hip_vjoin(gem_ctx *c, const uint32_t *vertex_order, uint64_t cap, gem_join_result *out, uint64_t *level_rows) {
  return guard(c, [&]() {
    if (!out) throw Error{GEM_E_INVALID, "null result"};
    run_vjoin(c, vertex_order, cap, out, level_rows);
  });
}

int gem_hip_order_candidates(gem_ctx *c, uint32_t max_orders, uint64_t seed, uint32_t *orders, int32_t *feats,
                             uint32_t *n) {
  return guard(c, [&]() {
    if (!n) throw Error{GEM_E_INVALID, "null count"};
    run_order_candidates(c, max_orders, seed, orders, feats, n);
  });
}

int gem_hip_match_batch_learned(gem_ctx *c, const gem_query_view *queries, uint32_t nq, gem_mlp *mlp,
                                uint32_t max_orders, const gem_batch_opts *opts, gem_join_result *results,
                                uint64_t *level_rows, uint32_t *chosen) {
  return guard(c, [&]() {
    if (nq && (!queries || !results)) throw Error{GEM_E_INVALID, "null argument"};
    run_match_batch_learned(c, queries, nq, mlp, max_orders, opts, results, level_rows, chosen);
  });
}

int gem_hip_match_relaxed(gem_ctx *c, const gem_query_view *q, uint32_t threshold, gem_mlp *mlp,
                          uint32_t max_orders, const gem_batch_opts *opts, uint32_t max_rq, uint64_t *masks,
                          gem_join_result *results, uint32_t *num_rq) {
  return guard(c, [&]() {
    if (!q || !masks || !results || !num_rq || !max_rq) throw Error{GEM_E_INVALID, "null argument"};
    gemh::QueryInfo Q;
    std::string why;
    int st = gemh::analyse_query(q, &Q, &why);
    if (st) throw Error{st, why};
    std::vector<gemh::EdgeSet> sets;
    st = gemh::relaxed_edge_sets(Q, threshold, max_rq, &sets);
    if (st) throw Error{st, "more than max_rq relaxed queries"};
    run_match_relaxed(c, q, threshold, sets, mlp, max_orders, opts, results);
    gemh::pack_edge_sets(sets, masks);
    *num_rq = (uint32_t)sets.size();
  });
}

int gem_hip_phase_times(const gem_ctx *c, gem_phase_times *out) {
  if (!c || !out) return GEM_E_INVALID;
  *out = c->times;
  return GEM_OK;
}

int gem_hip_match(gem_ctx *c, const gem_query_view *q, uint64_t cap, gem_join_result *out, uint32_t *order_out) {
  return guard(c, [&]() {
    if (!out) throw Error{GEM_E_INVALID, "null result"};
    const std::vector<uint32_t> order = plain_prepare(c, q);
    const gemh::QueryInfo &Q = c->q.Q;
    if (order_out) std::copy(order.begin(), order.end(), order_out);
    JoinOut jo;
    run_join(c, order.data(), cap, &jo, nullptr, nullptr, 0);
    out->count = jo.count;
    out->saturated = jo.saturated ? 1 : 0;
    out->num_steps = Q.ne - 1;
    out->num_embeddings_written = 0;
    out->rows_processed = jo.rows_processed;
    out->work_limited = jo.work_limited ? 1 : 0;
    out->restarts = 0;
    out->work = jo.rows_processed;
  });
}

int gem_hip_export_embeddings(gem_ctx *c, const gem_query_view *q, const char *path, uint64_t max_rows,
                              uint64_t run_bytes, gem_export_result *out) {
  return guard(c, [&]() {
    if (!q || !path) throw Error{GEM_E_INVALID, "null argument"};
    const std::vector<uint32_t> order = plain_prepare(c, q);
    struct FileSink : EmbSink {
      gemh::EmbWriter w;
      FileSink(const char *p, uint32_t nv, uint64_t rb, uint64_t mx) : w(p, nv, rb, mx) {}
      bool put(const uint32_t *r, uint64_t n, uint32_t) override { return w.put(r, n); }
    } sink(path, c->q.Q.nv, run_bytes, max_rows);
    JoinOut jo;
    run_join_sink(c, order.data(), 0, &jo, nullptr, &sink);
    if (jo.work_limited) {
      sink.w.discard();
      throw Error{GEM_E_LIMIT, "work limit reached before every embedding was enumerated"};
    }
    if (jo.sink_stopped) {
      const int st = sink.w.status();
      sink.w.discard();
      throw Error{st ? st : GEM_E_IO, st == GEM_E_LIMIT ? "more than max_rows embeddings" : "writing embeddings failed"};
    }
    const int st = sink.w.finish(out);
    if (st) throw Error{st, "writing embeddings failed"};
  });
}

}  // extern "C"
