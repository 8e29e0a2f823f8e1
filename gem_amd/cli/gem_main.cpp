// Command-line front end with the reference's flags (in-core/src/main.cu:25-44,
// out-of-core/src/main.cu:412-460):
//   gem -q <query file | directory of Q_<id>.in> -d <data graph> [-g gpu]
//       [-t threshold (0..10, default 2)] [--cap N] [--no-fork] [--export FILE]
// It loads the data graph in the reference text format once, on the host, and
// runs each query through the C ABI: at threshold 0 the plain pipeline (phase
// lines and the result line of in-core/src/main.cu:147-236), above it every
// relaxed query of the query in one batched run (one "Relaxed Query ET <bits>:
// <count> matches" line each, main.cu:229-236).
//
// A directory runs its queries in id order, each in a child process of its own
// (out-of-core/src/main.cu:334-410): the parent never touches the GPU; the
// child opens it, uploads the data graph, matches and exits, and a child that
// fails or dies is reported ("[batch-fail] ...") while the batch continues.
// --no-fork runs the directory in this process with one upload instead.
// --export (one query file, threshold 0) also writes every embedding to FILE
// in canonical order (include/gem_hip.h, gem_hip_export_embeddings).
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>
#include <dirent.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "gem_hip.h"

namespace {

void usage() {
  std::fprintf(stderr,
               "usage: gem -q <query file|dir> -d <data graph> [-g gpu] [-t threshold 0..10 (=2)] [--cap N] "
               "[--no-fork] [--export FILE]\n");
}

bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_file(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// Query id of a file stem: an optional 'Q' or 'q', an optional '_' or '-',
// then decimal digits to the end (out-of-core/src/main.cu:50-72).
bool parse_query_id(const std::string &stem, uint32_t *id) {
  size_t p = 0;
  if (p < stem.size() && (stem[p] == 'Q' || stem[p] == 'q')) ++p;
  if (p < stem.size() && (stem[p] == '_' || stem[p] == '-')) ++p;
  if (p >= stem.size()) return false;
  uint64_t v = 0;
  for (size_t i = p; i < stem.size(); ++i) {
    if (stem[i] < '0' || stem[i] > '9') return false;
    v = v * 10 + (uint64_t)(stem[i] - '0');
    if (v > 0xFFFFFFFFull) return false;
  }
  *id = (uint32_t)v;
  return true;
}

struct Job {
  uint32_t id;
  std::string path, name;
};

// Every regular *.in file of the directory, by id; a *.in whose stem is not a
// query id, or two files with one id, is an error (out-of-core/src/main.cu:74-124).
bool directory_jobs(const std::string &dir, std::vector<Job> *jobs, std::string *err) {
  DIR *d = opendir(dir.c_str());
  if (!d) {
    *err = "cannot open query directory " + dir;
    return false;
  }
  std::set<uint32_t> seen;
  bool ok = true;
  while (dirent *e = readdir(d)) {
    const std::string n = e->d_name;
    if (n.size() < 4 || n.compare(n.size() - 3, 3, ".in") != 0) continue;
    const std::string path = dir + "/" + n;
    if (!is_file(path)) continue;
    uint32_t id = 0;
    if (!parse_query_id(n.substr(0, n.size() - 3), &id)) {
      *err = "query file does not match expected pattern Q_<id>.in: " + path;
      ok = false;
      break;
    }
    if (!seen.insert(id).second) {
      *err = "duplicate query id " + std::to_string(id) + " in directory: " + dir;
      ok = false;
      break;
    }
    jobs->push_back({id, path, n});
  }
  closedir(d);
  std::sort(jobs->begin(), jobs->end(), [](const Job &a, const Job &b) { return a.id < b.id; });
  return ok;
}

// The reference prints edge tags as std::bitset<32> (in-core/lib/utils/include/
// defs.h:23-39): edge i is character 31 - i.  Wider queries print num_e bits.
std::string edge_bits(const uint64_t *mask, uint32_t ne) {
  const uint32_t width = std::max<uint32_t>(32, ne);
  std::string s(width, '0');
  for (uint32_t i = 0; i < ne; ++i)
    if ((mask[i / 64] >> (i % 64)) & 1) s[width - 1 - i] = '1';
  return s;
}

long long us_since(std::chrono::steady_clock::time_point t0) {
  return (long long)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0)
      .count();
}

std::string g_export;  // --export FILE

int run_query(gem_ctx *ctx, const std::string &path, uint32_t threshold, uint64_t cap) {
  uint32_t nv = 0, ne = 0;
  std::vector<uint32_t> lab(GEM_MAX_QV), edges(2 * GEM_MAX_QE);
  int st = gem_query_load(path.c_str(), GEM_MAX_QV, GEM_MAX_QE, &nv, lab.data(), &ne, edges.data());
  if (st) {
    std::fprintf(stderr, "%s: cannot load query (status %d)\n", path.c_str(), st);
    return st;
  }
  gem_query_view q{nv, ne, lab.data(), edges.data()};
  const auto t0 = std::chrono::steady_clock::now();
  if (threshold == 0) {
    gem_join_result r{};
    st = gem_hip_match(ctx, &q, cap, &r, nullptr);
    if (st) {
      std::fprintf(stderr, "%s: %s (status %d)\n", path.c_str(), gem_hip_last_error(ctx), st);
      return st;
    }
    const long long total = us_since(t0);
    gem_phase_times pt{};
    gem_hip_phase_times(ctx, &pt);
    std::printf("Filter time: %lld us\n", (long long)(pt.filter_ms * 1000));
    std::printf("Build Hash Table time: %lld us\n", (long long)(pt.index_ms * 1000));
    std::printf("construct time: %lld us\n", (long long)(pt.tables_ms * 1000));
    std::printf("Join Time: %lld us\n", (long long)(pt.join_ms * 1000));
    std::printf("Total time: %lld us\n", total);
    std::vector<uint64_t> all(GEM_RQ_WORDS, 0);
    for (uint32_t i = 0; i < ne; ++i) all[i / 64] |= 1ull << (i % 64);
    std::printf("\n=== Final Results ===\nRelaxed Query ET %s: %llu matches%s\n\n", edge_bits(all.data(), ne).c_str(),
                (unsigned long long)r.count, r.saturated ? " (capped)" : "");
    if (!g_export.empty()) {
      gem_export_result er{};
      st = gem_hip_export_embeddings(ctx, &q, g_export.c_str(), 0, 0, &er);
      if (st) {
        std::fprintf(stderr, "%s: export: %s (status %d)\n", path.c_str(), gem_hip_last_error(ctx), st);
        return st;
      }
      std::printf("Embeddings written: %llu to %s (digest %016llx)\n", (unsigned long long)er.rows,
                  g_export.c_str(), (unsigned long long)er.digest);
    }
    return 0;
  }
  uint32_t nrq = 0;
  st = gem_plan_relaxed_queries(&q, threshold, 0, nullptr, &nrq);
  if (st) {
    std::fprintf(stderr, "%s: relaxed queries at threshold %u: status %d\n", path.c_str(), threshold, st);
    return st;
  }
  std::vector<uint64_t> masks((size_t)nrq * GEM_RQ_WORDS);
  std::vector<gem_join_result> res(nrq);
  gem_batch_opts o{};
  o.cap = cap;
  uint32_t got = 0;
  st = gem_hip_match_relaxed(ctx, &q, threshold, nullptr, 1, &o, nrq, masks.data(), res.data(), &got);
  if (st) {
    std::fprintf(stderr, "%s: %s (status %d)\n", path.c_str(), gem_hip_last_error(ctx), st);
    return st;
  }
  const long long total = us_since(t0);
  std::printf("Relaxed queries: %u (threshold %u)\n", got, threshold);
  std::printf("Join Time: %lld us\n", total);
  std::printf("Join Time: %lld ms\n", total / 1000);
  std::printf("\n=== Final Results ===\n");
  for (uint32_t i = 0; i < got; ++i)
    std::printf("Relaxed Query ET %s: %llu matches%s\n", edge_bits(&masks[(size_t)i * GEM_RQ_WORDS], ne).c_str(),
                (unsigned long long)res[i].count, res[i].saturated ? " (capped)" : "");
  std::printf("\n");
  return 0;
}

// Opens the GPU and uploads the host graph.  nullptr on failure (reported).
gem_ctx *open_device(int gpu, gem_graph *g) {
  gem_ctx *ctx = nullptr;
  int st = gem_hip_create(gpu, &ctx);
  if (st) {
    std::fprintf(stderr, "cannot open GPU %d (status %d)\n", gpu, st);
    return nullptr;
  }
  gem_csr_view view{gem_graph_num_vertices(g), gem_graph_num_slots(g), gem_graph_offsets(g), gem_graph_neighbors(g),
                    gem_graph_labels(g)};
  st = gem_hip_load_data_csr(ctx, &view);
  if (st) {
    std::fprintf(stderr, "upload failed: %s\n", gem_hip_last_error(ctx));
    gem_hip_destroy(ctx);
    return nullptr;
  }
  return ctx;
}

int run_in_process(const std::vector<Job> &jobs, gem_graph *g, int gpu, uint32_t threshold, uint64_t cap,
                   bool batch) {
  gem_ctx *ctx = open_device(gpu, g);
  if (!ctx) return 1;
  size_t failed = 0;
  for (const Job &j : jobs) {
    if (batch) std::printf("%u\n", j.id);
    if (run_query(ctx, j.path, threshold, cap)) ++failed;
  }
  if (batch) std::printf("=== Batch Summary ===\nsuccess=%zu failed=%zu\n", jobs.size() - failed, failed);
  gem_hip_destroy(ctx);
  return batch ? 0 : (failed ? 1 : 0);
}

int run_forked(const std::vector<Job> &jobs, gem_graph *g, int gpu, uint32_t threshold, uint64_t cap) {
  struct Failure {
    const Job *job;
    std::string detail;
  };
  std::vector<Failure> failures;
  size_t ok = 0;
  for (const Job &j : jobs) {
    std::printf("%u\n", j.id);
    std::fflush(nullptr);
    const pid_t pid = fork();
    if (pid < 0) {
      std::fprintf(stderr, "ERROR: fork() failed for query %u: %s\n", j.id, std::strerror(errno));
      return 1;
    }
    if (pid == 0) {
      int rc = 1;
      if (gem_ctx *ctx = open_device(gpu, g)) {
        rc = run_query(ctx, j.path, threshold, cap) ? 1 : 0;
        gem_hip_destroy(ctx);
      }
      std::fflush(nullptr);
      _exit(rc);
    }
    int status = 0;
    pid_t w;
    do {
      w = waitpid(pid, &status, 0);
    } while (w < 0 && errno == EINTR);
    if (w < 0) {
      std::fprintf(stderr, "ERROR: waitpid() failed for query %u: %s\n", j.id, std::strerror(errno));
      return 1;
    }
    if (WIFEXITED(status) && WEXITSTATUS(status) == 0) {
      ++ok;
      continue;
    }
    std::string detail = WIFEXITED(status)     ? "exit=" + std::to_string(WEXITSTATUS(status))
                         : WIFSIGNALED(status) ? "signal=" + std::to_string(WTERMSIG(status))
                                               : "status=" + std::to_string(status);
    std::printf("[batch-fail] query %u file=%s %s continuing\n", j.id, j.name.c_str(), detail.c_str());
    failures.push_back({&j, detail});
  }
  std::printf("=== Batch Summary ===\nsuccess=%zu failed=%zu\n", ok, failures.size());
  for (const Failure &f : failures)
    std::printf("[failed] query %u file=%s %s\n", f.job->id, f.job->name.c_str(), f.detail.c_str());
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::string qpath, dpath;
  int gpu = 0;
  long threshold = 2;
  uint64_t cap = 0;
  bool fork_each = true;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
    const char *v = nullptr;
    if ((a == "-q" || a == "--query") && (v = next())) qpath = v;
    else if ((a == "-d" || a == "--data") && (v = next())) dpath = v;
    else if ((a == "-g" || a == "--gpu") && (v = next())) gpu = std::atoi(v);
    else if ((a == "-t" || a == "--threshold") && (v = next())) threshold = std::atol(v);
    else if (a == "--cap" && (v = next())) cap = std::strtoull(v, nullptr, 10);
    else if (a == "--eviction" && (v = next())) {
      // accepted and ignored, as the reference's command line does
      // (out-of-core/src/main.cu:128-130, 154: always LRU)
    }
    else if (a == "--no-fork") fork_each = false;
    else if (a == "--export" && (v = next())) g_export = v;
    else {
      usage();
      return 2;
    }
  }
  if (qpath.empty() || dpath.empty()) {
    usage();
    return 2;
  }
  if (threshold < 0 || threshold > 10) {
    std::fprintf(stderr, "--threshold must be in [0, 10]\n");
    return 2;
  }
  std::vector<Job> jobs;
  const bool batch = is_dir(qpath);
  if (!g_export.empty() && (batch || threshold != 0)) {
    std::fprintf(stderr, "--export takes one query file at --threshold 0\n");
    return 2;
  }
  if (batch) {
    std::string err;
    if (!directory_jobs(qpath, &jobs, &err)) {
      std::fprintf(stderr, "ERROR: %s\n", err.c_str());
      return 1;
    }
  } else if (is_file(qpath)) {
    const size_t slash = qpath.find_last_of('/');
    const std::string name = slash == std::string::npos ? qpath : qpath.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    uint32_t id = 0;
    if (!parse_query_id(dot == std::string::npos ? name : name.substr(0, dot), &id)) id = 0;
    jobs.push_back({id, qpath, name});
  } else {
    std::fprintf(stderr, "ERROR: Query path is not a file or directory: %s\n", qpath.c_str());
    return 1;
  }
  if (jobs.empty()) {
    std::fprintf(stderr, "ERROR: No query graphs found in: %s\n", qpath.c_str());
    return 1;
  }
  gem_graph *g = nullptr;
  int st = gem_graph_load(dpath.c_str(), &g);
  if (st) {
    std::fprintf(stderr, "cannot load data graph %s (status %d)\n", dpath.c_str(), st);
    return 1;
  }
  const int rc = batch && fork_each ? run_forked(jobs, g, gpu, (uint32_t)threshold, cap)
                                    : run_in_process(jobs, g, gpu, (uint32_t)threshold, cap, batch);
  gem_graph_free(g);
  return rc;
}
