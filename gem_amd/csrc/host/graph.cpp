// Host graph substrate: deterministic generators (Erdos-Renyi, R-MAT, uniform
// and Zipf labels), the reference text format (reader and writer) and
// random-walk query extraction.  Exposed through include/gem_hip.h.
//
// The text format follows the reference reader
// (in-core/lib/io/src/io.cu:45-220): a "t N M" header, one "v id label degree"
// line per vertex, one "e src dst" line per undirected edge.  The reference
// places neighbours in file order and requires them sorted for its binary
// searches (in-core/lib/join/src/join_trie_kernel.cu:148); this reader sorts
// each list, which equals the reference's layout whenever the file is valid
// input for it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <unordered_set>

#include "gem_hip.h"
#include "host_graph.h"

namespace gemh {

bool validate_csr(const gem_graph &g, std::string *why) {
  auto fail = [&](const std::string &m) {
    if (why) *why = m;
    return false;
  };
  if (g.off.size() != (size_t)g.n + 1 || g.lab.size() != g.n) return fail("array sizes");
  if (g.off[0] != 0 || g.off[g.n] != g.nbr.size()) return fail("offsets bounds");
  for (uint32_t v = 0; v < g.n; ++v) {
    if (g.off[v + 1] < g.off[v]) return fail("offsets not monotone");
    for (uint64_t i = g.off[v]; i < g.off[v + 1]; ++i) {
      uint32_t w = g.nbr[i];
      if (w >= g.n) return fail("neighbour out of range");
      if (w == v) return fail("self loop at " + std::to_string(v));
      if (i > g.off[v] && g.nbr[i - 1] >= w) return fail("adjacency of " + std::to_string(v) + " not strictly ascending");
      // symmetric: v must appear in adj(w)
      const uint32_t *b = g.nbr.data() + g.off[w], *e = g.nbr.data() + g.off[w + 1];
      if (!std::binary_search(b, e, v)) return fail("asymmetric edge");
    }
  }
  return true;
}

static void assign_labels(gem_graph &g, uint32_t num_labels, uint32_t dist, uint64_t seed) {
  g.lab.resize(g.n);
  std::vector<double> cdf(num_labels);
  double tot = 0;
  for (uint32_t k = 0; k < num_labels; ++k) {
    tot += (dist == 1) ? 1.0 / (double)(k + 1) : 1.0;
    cdf[k] = tot;
  }
  uint64_t lseed = splitmix64(seed ^ 0x6c6162656cull);  // "label"
  parallel_chunks(g.n, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; ++v) {
      Rng r(splitmix64(lseed + v));
      double u = r.uniform() * tot;
      uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
      g.lab[v] = k < num_labels ? k : num_labels - 1;
    }
  });
}

}  // namespace gemh

using namespace gemh;

extern "C" {

int gem_graph_generate_er(uint32_t n, uint64_t m, uint32_t num_labels, uint64_t seed, gem_graph **out) {
  if (!out || n < 2 || num_labels == 0) return GEM_E_INVALID;
  if (m > (uint64_t)n * (n - 1) / 2) return GEM_E_INVALID;
  // G(n, m): draw ordered pairs from one stream, keep the first m distinct
  // unordered pairs.
  std::unordered_set<uint64_t> seen;
  seen.reserve(m * 2);
  std::vector<uint64_t> pairs;
  pairs.reserve(m);
  Rng r(splitmix64(seed));
  while (pairs.size() < m) {
    uint32_t a = (uint32_t)r.below(n), b = (uint32_t)r.below(n);
    if (a == b) continue;
    uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
    if (seen.insert(key).second) pairs.push_back(key);
  }
  auto *g = new gem_graph();
  build_symmetric_csr(n, pairs.size(), [&](uint64_t i, uint32_t *s, uint32_t *d) {
    *s = (uint32_t)(pairs[i] >> 32);
    *d = (uint32_t)pairs[i];
    return true;
  }, *g);
  assign_labels(*g, num_labels, 0, seed);
  *out = g;
  return GEM_OK;
}

int gem_graph_generate_rmat(uint32_t scale, double avg_degree, double a, double b, double c,
                            uint32_t num_labels, uint32_t label_dist, uint64_t seed, gem_graph **out) {
  if (!out || scale == 0 || scale > 31 || num_labels == 0 || avg_degree <= 0) return GEM_E_INVALID;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return GEM_E_INVALID;
  const uint32_t n = 1u << scale;
  const uint64_t samples = (uint64_t)std::llround((double)n * avg_degree / 2.0);
  const double ab = a + b, abc = a + b + c;
  const uint64_t eseed = splitmix64(seed ^ 0x65646765ull);  // "edge"
  auto *g = new gem_graph();
  // Sample i: one quadrant choice per level, most significant bit first.
  build_symmetric_csr(n, samples, [&](uint64_t i, uint32_t *s, uint32_t *d) {
    Rng r(splitmix64(eseed + i));
    uint32_t src = 0, dst = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      double u = r.uniform();
      uint32_t sb = 0, db = 0;
      if (u < a) {
      } else if (u < ab) {
        db = 1;
      } else if (u < abc) {
        sb = 1;
      } else {
        sb = 1;
        db = 1;
      }
      src = (src << 1) | sb;
      dst = (dst << 1) | db;
    }
    *s = src;
    *d = dst;
    return true;
  }, *g);
  assign_labels(*g, num_labels, label_dist, seed);
  *out = g;
  return GEM_OK;
}

int gem_graph_from_csr(uint32_t n, const uint64_t *offsets, const uint32_t *neighbors,
                       const uint32_t *labels, gem_graph **out) {
  if (!out || !offsets || !labels || (offsets[n] && !neighbors)) return GEM_E_INVALID;
  auto *g = new gem_graph();
  g->n = n;
  g->off.assign(offsets, offsets + n + 1);
  g->nbr.assign(neighbors, neighbors + offsets[n]);
  g->lab.assign(labels, labels + n);
  std::string why;
  if (!validate_csr(*g, &why)) {
    delete g;
    return GEM_E_INVALID;
  }
  *out = g;
  return GEM_OK;
}

int gem_graph_from_csr32(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, const uint32_t *labels,
                         gem_graph **out) {
  if (!out || !offsets || !labels || (offsets[n] && !neighbors)) return GEM_E_INVALID;
  *out = nullptr;
  if (offsets[0] != 0) return GEM_E_INVALID;
  for (uint32_t v = 0; v < n; ++v)
    if (offsets[v + 1] < offsets[v]) return GEM_E_INVALID;
  auto *g = new gem_graph();
  g->n = n;
  g->off.assign(offsets, offsets + n + 1);
  g->nbr.assign(neighbors, neighbors + offsets[n]);
  g->lab.assign(labels, labels + n);
  // the reference's loader appends neighbours in edge-line order
  gemh::parallel_chunks(n, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; ++v) std::sort(g->nbr.begin() + g->off[v], g->nbr.begin() + g->off[v + 1]);
  });
  std::string why;
  if (!gemh::validate_csr(*g, &why)) {
    delete g;
    return GEM_E_INVALID;
  }
  *out = g;
  return GEM_OK;
}

// Minimal whitespace-separated token reader over a whole file.
namespace {
struct Tok {
  std::vector<char> buf;
  size_t p = 0;
  bool open(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    size_t rd = sz > 0 ? std::fread(buf.data(), 1, buf.size(), f) : 0;
    std::fclose(f);
    return rd == buf.size();
  }
  void skip() {
    while (p < buf.size() && (buf[p] == ' ' || buf[p] == '\t' || buf[p] == '\n' || buf[p] == '\r')) ++p;
  }
  int chr() {
    skip();
    return p < buf.size() ? buf[p++] : -1;
  }
  bool u64(uint64_t *v) {
    skip();
    if (p >= buf.size() || buf[p] < '0' || buf[p] > '9') return false;
    uint64_t x = 0;
    while (p < buf.size() && buf[p] >= '0' && buf[p] <= '9') x = x * 10 + (uint64_t)(buf[p++] - '0');
    *v = x;
    return true;
  }
};

struct TextGraph {
  uint64_t n = 0, m = 0;
  std::vector<uint32_t> lab, deg;
  std::vector<std::pair<uint32_t, uint32_t>> edges;
};

int read_text(const char *path, TextGraph &t) {
  Tok k;
  if (!path || !k.open(path)) return GEM_E_IO;
  if (k.chr() != 't' || !k.u64(&t.n) || !k.u64(&t.m)) return GEM_E_INVALID;
  if (t.n > 0xFFFFFFFFull) return GEM_E_LIMIT;
  t.lab.assign(t.n, 0);
  t.deg.assign(t.n, 0);
  std::vector<uint8_t> seen(t.n, 0);
  t.edges.reserve(t.m);
  for (;;) {
    int c = k.chr();
    if (c < 0) break;
    if (c == 'v') {
      uint64_t id, l, d;
      if (!k.u64(&id) || !k.u64(&l) || !k.u64(&d) || id >= t.n || l > 0xFFFFFFFFull) return GEM_E_INVALID;
      t.lab[id] = (uint32_t)l;
      t.deg[id] = (uint32_t)d;
      seen[id] = 1;
    } else if (c == 'e') {
      uint64_t s, d;
      if (!k.u64(&s) || !k.u64(&d) || s >= t.n || d >= t.n) return GEM_E_INVALID;
      t.edges.emplace_back((uint32_t)s, (uint32_t)d);
    } else {
      return GEM_E_INVALID;
    }
  }
  for (uint64_t v = 0; v < t.n; ++v)
    if (!seen[v]) return GEM_E_INVALID;
  if (t.edges.size() != t.m) return GEM_E_INVALID;
  return GEM_OK;
}
}  // namespace

// Binary CSR cache: a 64-byte header ("GEMCSR01", n, slots, label count
// and the FNV-1a 64 digest of the three arrays), then offsets[n + 1] (u64),
// neighbors[slots] (u32) and labels[n] (u32), little-endian as in memory.
namespace {
constexpr char BIN_MAGIC[8] = {'G', 'E', 'M', 'C', 'S', 'R', '0', '1'};
struct BinHeader {
  char magic[8];
  uint64_t n, slots, reserved, digest;
  uint64_t pad[3];
};
static_assert(sizeof(BinHeader) == 64, "header is 64 bytes");

uint64_t fnv1a(uint64_t h, const void *p, size_t bytes) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

// Digest over fixed 1 MiB pieces, combined in order: the same value however
// many threads hash the pieces.
uint64_t csr_digest(const gem_graph &g) {
  struct Span {
    const void *p;
    size_t bytes;
  };
  const Span arrays[3] = {{g.off.data(), g.off.size() * 8}, {g.nbr.data(), g.nbr.size() * 4},
                          {g.lab.data(), g.lab.size() * 4}};
  std::vector<Span> pieces;
  constexpr size_t PIECE = 1u << 20;
  for (const Span &a : arrays)
    for (size_t o = 0; o < a.bytes; o += PIECE)
      pieces.push_back({static_cast<const char *>(a.p) + o, std::min(PIECE, a.bytes - o)});
  std::vector<uint64_t> h(pieces.size());
  gemh::parallel_chunks(pieces.size(), [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) h[i] = fnv1a(0xcbf29ce484222325ull, pieces[i].p, pieces[i].bytes);
  });
  return fnv1a(0xcbf29ce484222325ull, h.data(), h.size() * 8);
}

bool read_all(FILE *f, void *p, size_t bytes) {
  char *c = static_cast<char *>(p);
  while (bytes) {
    const size_t k = std::fread(c, 1, std::min<size_t>(bytes, 1u << 30), f);
    if (k == 0) return false;
    c += k;
    bytes -= k;
  }
  return true;
}

int load_binary(FILE *f, gem_graph **out) {
  BinHeader h;
  if (!read_all(f, &h, sizeof h) || std::memcmp(h.magic, BIN_MAGIC, 8) != 0) return GEM_E_INVALID;
  if (h.n > 0xFFFFFFFFull || h.slots > (1ull << 40)) return GEM_E_LIMIT;
  auto *g = new gem_graph();
  g->n = (uint32_t)h.n;
  g->off.resize(h.n + 1);
  g->nbr.resize(h.slots);
  g->lab.resize(h.n);
  const bool ok = read_all(f, g->off.data(), g->off.size() * 8) && read_all(f, g->nbr.data(), g->nbr.size() * 4) &&
                  read_all(f, g->lab.data(), g->lab.size() * 4) && std::fgetc(f) == EOF;
  if (!ok || g->off[0] != 0 || g->off[h.n] != h.slots || csr_digest(*g) != h.digest) {
    delete g;
    return GEM_E_INVALID;
  }
  // the digest guards the bytes; these checks guard the indexing
  bool bad = false;
  for (uint32_t v = 0; v < g->n && !bad; ++v) bad = g->off[v + 1] < g->off[v];
  std::vector<uint8_t> out_of_range(64, 0);
  gemh::parallel_chunks(g->nbr.size(), [&](uint64_t b, uint64_t e) {
    uint8_t any = 0;
    for (uint64_t i = b; i < e; ++i) any |= g->nbr[i] >= g->n;
    if (any) __atomic_store_n(&out_of_range[0], 1, __ATOMIC_RELAXED);
  });
  if (bad || out_of_range[0]) {
    delete g;
    return GEM_E_INVALID;
  }
  *out = g;
  return GEM_OK;
}
}  // namespace

int gem_graph_load(const char *path, gem_graph **out) {
  if (!out) return GEM_E_INVALID;
  if (path) {
    // a binary CSR cache is recognised by its magic
    if (FILE *f = std::fopen(path, "rb")) {
      char m[8] = {0};
      const bool bin = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, BIN_MAGIC, 8) == 0;
      int st = GEM_OK;
      if (bin) {
        std::rewind(f);
        st = load_binary(f, out);
      }
      std::fclose(f);
      if (bin) return st;
    }
  }
  TextGraph t;
  int st = read_text(path, t);
  if (st) return st;
  auto *g = new gem_graph();
  g->n = (uint32_t)t.n;
  build_symmetric_csr(g->n, t.edges.size(), [&](uint64_t i, uint32_t *s, uint32_t *d) {
    *s = t.edges[i].first;
    *d = t.edges[i].second;
    return true;
  }, *g);
  g->lab = t.lab;
  // The reference builds offsets from the declared degrees (io.cu:172); a file
  // whose degrees disagree with its edges (or has loops / repeated edges) is
  // rejected rather than silently mis-laid.
  for (uint32_t v = 0; v < g->n; ++v)
    if (g->off[v + 1] - g->off[v] != t.deg[v]) {
      delete g;
      return GEM_E_INVALID;
    }
  if (g->off[g->n] != 2 * t.m) {
    delete g;
    return GEM_E_INVALID;
  }
  *out = g;
  return GEM_OK;
}

int gem_graph_save(const gem_graph *g, const char *path) {
  if (!g || !path) return GEM_E_INVALID;
  FILE *f = std::fopen(path, "wb");
  if (!f) return GEM_E_IO;
  std::fprintf(f, "t %u %llu\n", g->n, (unsigned long long)(g->off[g->n] / 2));
  for (uint32_t v = 0; v < g->n; ++v)
    std::fprintf(f, "v %u %u %llu\n", v, g->lab[v], (unsigned long long)(g->off[v + 1] - g->off[v]));
  // edges (src < dst) in lexicographic order: the layout for which the
  // reference's file-order adjacency is sorted
  for (uint32_t v = 0; v < g->n; ++v)
    for (uint64_t i = g->off[v]; i < g->off[v + 1]; ++i)
      if (g->nbr[i] > v) std::fprintf(f, "e %u %u\n", v, g->nbr[i]);
  int bad = std::ferror(f);
  std::fclose(f);
  return bad ? GEM_E_IO : GEM_OK;
}

int gem_graph_save_binary(const gem_graph *g, const char *path) {
  if (!g || !path) return GEM_E_INVALID;
  BinHeader h{};
  std::memcpy(h.magic, BIN_MAGIC, 8);
  h.n = g->n;
  h.slots = g->off[g->n];
  h.digest = csr_digest(*g);
  FILE *f = std::fopen(path, "wb");
  if (!f) return GEM_E_IO;
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  ok = ok && std::fwrite(g->off.data(), 8, g->off.size(), f) == g->off.size();
  ok = ok && std::fwrite(g->nbr.data(), 4, g->nbr.size(), f) == g->nbr.size();
  ok = ok && std::fwrite(g->lab.data(), 4, g->lab.size(), f) == g->lab.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GEM_OK : GEM_E_IO;
}

uint64_t gem_graph_digest(const gem_graph *g) { return g ? csr_digest(*g) : 0; }

uint32_t gem_graph_num_vertices(const gem_graph *g) { return g ? g->n : 0; }
uint64_t gem_graph_num_slots(const gem_graph *g) { return g ? g->off[g->n] : 0; }
const uint64_t *gem_graph_offsets(const gem_graph *g) { return g ? g->off.data() : nullptr; }
const uint32_t *gem_graph_neighbors(const gem_graph *g) { return g ? g->nbr.data() : nullptr; }
const uint32_t *gem_graph_labels(const gem_graph *g) { return g ? g->lab.data() : nullptr; }
void gem_graph_free(gem_graph *g) { delete g; }

int gem_query_random_walk(const gem_graph *g, uint32_t k, uint64_t seed, uint32_t *labels_out,
                          uint32_t *edges_out, uint32_t max_e, uint32_t *num_e_out) {
  if (!g || k < 2 || k > GEM_MAX_QV || !labels_out || !edges_out || !num_e_out) return GEM_E_INVALID;
  if (g->n < k) return GEM_E_INVALID;
  Rng r(splitmix64(seed ^ 0x7175657279ull));  // "query"
  std::vector<uint32_t> vis;
  for (int attempt = 0; attempt < 10000; ++attempt) {
    uint32_t start = (uint32_t)r.below(g->n);
    if (g->off[start + 1] == g->off[start]) continue;
    vis.assign(1, start);
    uint32_t cur = start;
    for (uint32_t step = 0; step < 100u * k && vis.size() < k; ++step) {
      uint64_t d = g->off[cur + 1] - g->off[cur];
      uint32_t nxt = g->nbr[g->off[cur] + r.below(d)];
      if (std::find(vis.begin(), vis.end(), nxt) == vis.end()) vis.push_back(nxt);
      cur = nxt;
    }
    if (vis.size() == k) break;
  }
  if (vis.size() != k) return GEM_E_QUERY;
  uint32_t ne = 0;
  for (uint32_t i = 0; i < k; ++i) {
    labels_out[i] = g->lab[vis[i]];
    const uint32_t *b = g->nbr.data() + g->off[vis[i]], *e = g->nbr.data() + g->off[vis[i] + 1];
    for (uint32_t j = i + 1; j < k; ++j)
      if (std::binary_search(b, e, vis[j])) {
        if (ne >= max_e) return GEM_E_LIMIT;
        edges_out[2 * ne] = i;
        edges_out[2 * ne + 1] = j;
        ++ne;
      }
  }
  *num_e_out = ne;
  return GEM_OK;
}

int gem_query_load(const char *path, uint32_t max_v, uint32_t max_e, uint32_t *num_v, uint32_t *labels,
                   uint32_t *num_e, uint32_t *edges) {
  if (!num_v || !labels || !num_e || !edges) return GEM_E_INVALID;
  TextGraph t;
  int st = read_text(path, t);
  if (st) return st;
  if (t.n > max_v || t.m > max_e) return GEM_E_LIMIT;
  std::vector<uint32_t> deg(t.n, 0);
  for (size_t i = 0; i < t.edges.size(); ++i) {
    uint32_t s = t.edges[i].first, d = t.edges[i].second;
    if (s == d) return GEM_E_INVALID;
    edges[2 * i] = std::min(s, d);
    edges[2 * i + 1] = std::max(s, d);
    deg[s]++;
    deg[d]++;
  }
  for (uint64_t v = 0; v < t.n; ++v) {
    if (deg[v] != t.deg[v]) return GEM_E_INVALID;
    labels[v] = t.lab[v];
  }
  *num_v = (uint32_t)t.n;
  *num_e = (uint32_t)t.m;
  return GEM_OK;
}

int gem_query_save(const gem_query_view *q, const char *path) {
  if (!q || !path) return GEM_E_INVALID;
  FILE *f = std::fopen(path, "wb");
  if (!f) return GEM_E_IO;
  std::vector<uint32_t> deg(q->num_v, 0);
  for (uint32_t i = 0; i < q->num_e; ++i) {
    deg[q->edges[2 * i]]++;
    deg[q->edges[2 * i + 1]]++;
  }
  std::fprintf(f, "t %u %u\n", q->num_v, q->num_e);
  for (uint32_t v = 0; v < q->num_v; ++v) std::fprintf(f, "v %u %u %u\n", v, q->labels[v], deg[v]);
  for (uint32_t i = 0; i < q->num_e; ++i) std::fprintf(f, "e %u %u\n", q->edges[2 * i], q->edges[2 * i + 1]);
  int bad = std::ferror(f);
  std::fclose(f);
  return bad ? GEM_E_IO : GEM_OK;
}

}  // extern "C"
