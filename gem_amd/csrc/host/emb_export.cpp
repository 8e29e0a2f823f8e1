// Canonical embedding files (see emb_export.h).
#include "emb_export.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <queue>

#include "host_graph.h"

namespace gemh {

namespace {

constexpr char EMB_MAGIC[8] = {'G', 'E', 'M', 'E', 'M', 'B', '0', '1'};
constexpr uint64_t FNV_OFFSET = 0xcbf29ce484222325ull;

struct EmbHeader {
  char magic[8];
  uint32_t nv, zero;
  uint64_t rows, digest;
};
static_assert(sizeof(EmbHeader) == 32, "header is 32 bytes");

uint64_t fnv1a(uint64_t h, const void *p, size_t bytes) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

bool less_row(const uint32_t *a, const uint32_t *b, uint32_t nv) {
  for (uint32_t c = 0; c < nv; ++c)
    if (a[c] != b[c]) return a[c] < b[c];
  return false;
}

// Buffered sequential reader of one sorted run.
struct RunReader {
  FILE *f = nullptr;
  std::vector<uint32_t> buf;
  size_t pos = 0, len = 0;  // in rows
  uint32_t nv = 0;
  bool open(const std::string &p, uint32_t nv_, size_t rows) {
    nv = nv_;
    f = std::fopen(p.c_str(), "rb");
    buf.resize(rows * nv);
    return f != nullptr;
  }
  const uint32_t *cur() {
    if (pos == len) {
      len = f ? std::fread(buf.data(), 4 * (size_t)nv, buf.size() / nv, f) : 0;
      pos = 0;
      if (len == 0) return nullptr;
    }
    return buf.data() + pos * nv;
  }
  void next() { ++pos; }
  ~RunReader() {
    if (f) std::fclose(f);
  }
};

}  // namespace

EmbWriter::EmbWriter(std::string path, uint32_t num_v, uint64_t run_bytes, uint64_t max_rows)
    : path_(std::move(path)), nv_(std::max<uint32_t>(num_v, 1)), max_rows_(max_rows) {
  if (run_bytes == 0) run_bytes = 256ull << 20;
  rows_per_run_ = std::min<uint64_t>(std::max<uint64_t>(1, run_bytes / (4ull * nv_)), 1ull << 31);
}

EmbWriter::~EmbWriter() {
  if (!finished_) discard();
}

void EmbWriter::discard() {
  for (const std::string &r : runs_) std::remove(r.c_str());
  runs_.clear();
  std::remove(path_.c_str());
  finished_ = true;
}

bool EmbWriter::put(const uint32_t *rows, uint64_t n) {
  if (status_) return false;
  if (max_rows_ && rows_ + n > max_rows_) {
    status_ = GEM_E_LIMIT;
    return false;
  }
  while (n) {
    const uint64_t room = rows_per_run_ - buf_.size() / nv_;
    const uint64_t k = std::min(room, n);
    buf_.insert(buf_.end(), rows, rows + k * nv_);
    rows += k * nv_;
    n -= k;
    rows_ += k;
    if (buf_.size() / nv_ == rows_per_run_ && flush_run()) return false;
  }
  return true;
}

void EmbWriter::sort_buffer() {
  const uint64_t n = buf_.size() / nv_;
  std::vector<uint32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  const uint32_t *b = buf_.data();
  const uint32_t nv = nv_;
  std::sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return less_row(b + (uint64_t)x * nv, b + (uint64_t)y * nv, nv); });
  std::vector<uint32_t> out(buf_.size());
  parallel_chunks(n, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) std::memcpy(&out[i * nv], b + (uint64_t)idx[i] * nv, 4ull * nv);
  });
  buf_.swap(out);
}

int EmbWriter::flush_run() {
  sort_buffer();
  std::string p = path_ + ".run" + std::to_string(runs_.size());
  FILE *f = std::fopen(p.c_str(), "wb");
  bool ok = f && std::fwrite(buf_.data(), 4, buf_.size(), f) == buf_.size();
  if (f) ok = (std::fclose(f) == 0) && ok;
  runs_.push_back(p);
  buf_.clear();
  if (!ok) status_ = GEM_E_IO;
  return status_;
}

int EmbWriter::finish(gem_export_result *out) {
  if (status_) return status_;
  FILE *f = std::fopen(path_.c_str(), "wb");
  if (!f) return status_ = GEM_E_IO;
  EmbHeader h{};
  std::memcpy(h.magic, EMB_MAGIC, 8);
  h.nv = nv_;
  h.rows = rows_;
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  uint64_t digest = FNV_OFFSET;
  if (runs_.empty()) {
    sort_buffer();
    digest = fnv1a(digest, buf_.data(), buf_.size() * 4);
    ok = ok && std::fwrite(buf_.data(), 4, buf_.size(), f) == buf_.size();
  } else {
    if (!buf_.empty() && flush_run()) {
      std::fclose(f);
      return status_;
    }
    // k-way merge of the sorted runs
    const size_t k = runs_.size();
    std::vector<RunReader> rd(k);
    const size_t rows_each = std::max<size_t>(1024, (64ull << 20) / (4ull * nv_ * k));
    for (size_t i = 0; i < k; ++i)
      if (!rd[i].open(runs_[i], nv_, rows_each)) ok = false;
    auto greater = [&](size_t a, size_t b) { return less_row(rd[b].cur(), rd[a].cur(), nv_); };
    std::priority_queue<size_t, std::vector<size_t>, decltype(greater)> heap(greater);
    for (size_t i = 0; ok && i < k; ++i)
      if (rd[i].cur()) heap.push(i);
    std::vector<uint32_t> wbuf;
    wbuf.reserve((size_t)nv_ * 65536);
    auto drain = [&]() {
      digest = fnv1a(digest, wbuf.data(), wbuf.size() * 4);
      ok = ok && std::fwrite(wbuf.data(), 4, wbuf.size(), f) == wbuf.size();
      wbuf.clear();
    };
    while (ok && !heap.empty()) {
      const size_t i = heap.top();
      heap.pop();
      const uint32_t *r = rd[i].cur();
      wbuf.insert(wbuf.end(), r, r + nv_);
      rd[i].next();
      if (rd[i].cur()) heap.push(i);
      if (wbuf.size() >= (size_t)nv_ * 65536) drain();
    }
    drain();
  }
  h.digest = digest;
  ok = ok && std::fseek(f, 0, SEEK_SET) == 0 && std::fwrite(&h, sizeof h, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  for (const std::string &r : runs_) std::remove(r.c_str());
  const uint32_t nruns = (uint32_t)std::max<size_t>(runs_.size(), 1);
  runs_.clear();
  if (!ok) {
    std::remove(path_.c_str());
    finished_ = true;
    return status_ = GEM_E_IO;
  }
  finished_ = true;
  if (out) {
    out->rows = rows_;
    out->digest = digest;
    out->runs = nruns;
    out->num_v = nv_;
  }
  return GEM_OK;
}

}  // namespace gemh

extern "C" int gem_export_rows(const uint32_t *rows, uint64_t n, uint32_t num_v, const char *path,
                               uint64_t run_bytes, gem_export_result *out) {
  if (!path || num_v == 0 || num_v > GEM_MAX_QV || (n && !rows)) return GEM_E_INVALID;
  gemh::EmbWriter w(path, num_v, run_bytes, 0);
  if (!w.put(rows, n)) return w.status();
  return w.finish(out);
}
