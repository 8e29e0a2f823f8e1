// Command-line front end with the reference's flags (in-core/src/main.cu:33-44):
//   gem -q <query file | directory of Q_<id>.in> -d <data graph> [-g gpu] [-t 0]
// It loads the graphs in the reference text format, runs the plain pipeline on
// the GPU through the C ABI and prints the reference's phase lines and result
// line (in-core/src/main.cu:147-236).  A query directory runs every Q_<id>.in
// in id order with one data-graph upload (out-of-core/src/main.cu:74-124,
// 334-410); a failing query is reported and the batch continues.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <bitset>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gem_hip.h"

static void usage() {
  std::fprintf(stderr,
               "usage: gem -q <query file|dir> -d <data graph> [-g gpu] [-t threshold(=0)] [--cap N]\n");
}

static bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

static std::vector<std::string> query_jobs(const std::string &dir) {
  std::vector<std::pair<long, std::string>> jobs;
  if (DIR *d = opendir(dir.c_str())) {
    while (dirent *e = readdir(d)) {
      std::string n = e->d_name;
      if (n.size() > 5 && n.rfind("Q_", 0) == 0 && n.substr(n.size() - 3) == ".in") {
        char *end = nullptr;
        long id = std::strtol(n.c_str() + 2, &end, 10);
        if (end && std::string(end) == ".in") jobs.push_back({id, dir + "/" + n});
      }
    }
    closedir(d);
  }
  std::sort(jobs.begin(), jobs.end());
  std::vector<std::string> out;
  for (auto &j : jobs) out.push_back(j.second);
  return out;
}

static int run_one(gem_ctx *ctx, const std::string &path, uint64_t cap) {
  uint32_t nv = 0, ne = 0;
  std::vector<uint32_t> lab(GEM_MAX_QV), edges(2 * GEM_MAX_QE);
  int st = gem_query_load(path.c_str(), GEM_MAX_QV, GEM_MAX_QE, &nv, lab.data(), &ne, edges.data());
  if (st) {
    std::fprintf(stderr, "%s: cannot load query (status %d)\n", path.c_str(), st);
    return st;
  }
  gem_query_view q{nv, ne, lab.data(), edges.data()};
  gem_join_result r{};
  auto t0 = std::chrono::steady_clock::now();
  st = gem_hip_match(ctx, &q, cap, &r, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  if (st) {
    std::fprintf(stderr, "%s: %s (status %d)\n", path.c_str(), gem_hip_last_error(ctx), st);
    return st;
  }
  gem_phase_times pt{};
  gem_hip_phase_times(ctx, &pt);
  std::printf("Filter time: %lld us\n", (long long)(pt.filter_ms * 1000));
  std::printf("Build Hash Table time: %lld us\n", (long long)(pt.index_ms * 1000));
  std::printf("construct time: %lld us\n", (long long)(pt.tables_ms * 1000));
  std::printf("Join Time: %lld us\n", (long long)(pt.join_ms * 1000));
  std::printf("Total time: %lld us\n",
              (long long)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
  std::string bits;
  for (int i = (int)ne - 1; i >= 0; --i) bits.push_back('1');
  std::printf("\n=== Final Results ===\nRelaxed Query ET %s: %llu matches%s\n\n", bits.c_str(),
              (unsigned long long)r.count, r.saturated ? " (capped)" : "");
  return 0;
}

int main(int argc, char **argv) {
  std::string qpath, dpath;
  int gpu = 0;
  long threshold = 0;
  uint64_t cap = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
    const char *v = nullptr;
    if ((a == "-q" || a == "--query") && (v = next())) qpath = v;
    else if ((a == "-d" || a == "--data") && (v = next())) dpath = v;
    else if ((a == "-g" || a == "--gpu") && (v = next())) gpu = std::atoi(v);
    else if ((a == "-t" || a == "--threshold") && (v = next())) threshold = std::atol(v);
    else if (a == "--cap" && (v = next())) cap = std::strtoull(v, nullptr, 10);
    else if (a == "--eviction" && (v = next())) { /* accepted for compatibility */ }
    else {
      usage();
      return 2;
    }
  }
  if (qpath.empty() || dpath.empty()) {
    usage();
    return 2;
  }
  if (threshold != 0) {
    std::fprintf(stderr, "only --threshold 0 (exact subgraph isomorphism) is supported\n");
    return 2;
  }
  gem_graph *g = nullptr;
  int st = gem_graph_load(dpath.c_str(), &g);
  if (st) {
    std::fprintf(stderr, "cannot load data graph %s (status %d)\n", dpath.c_str(), st);
    return 1;
  }
  gem_ctx *ctx = nullptr;
  st = gem_hip_create(gpu, &ctx);
  if (st) {
    std::fprintf(stderr, "cannot open GPU %d (status %d)\n", gpu, st);
    gem_graph_free(g);
    return 1;
  }
  gem_csr_view view{gem_graph_num_vertices(g), gem_graph_num_slots(g), gem_graph_offsets(g),
                    gem_graph_neighbors(g), gem_graph_labels(g)};
  st = gem_hip_load_data_csr(ctx, &view);
  if (st) {
    std::fprintf(stderr, "upload failed: %s\n", gem_hip_last_error(ctx));
    return 1;
  }
  int rc = 0;
  if (is_dir(qpath)) {
    auto jobs = query_jobs(qpath);
    size_t failed = 0;
    for (auto &j : jobs) {
      std::printf("=== %s ===\n", j.c_str());
      if (run_one(ctx, j, cap)) ++failed;
    }
    std::printf("Batch: %zu queries, %zu failed\n", jobs.size(), failed);
    rc = failed ? 1 : 0;
  } else {
    rc = run_one(ctx, qpath, cap) ? 1 : 0;
  }
  gem_hip_destroy(ctx);
  gem_graph_free(g);
  return rc;
}
