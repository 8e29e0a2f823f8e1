// Embedding export: rows of num_v vertex ids (column u = image of query vertex
// u) written to a file in canonical order -- ascending lexicographic over the
// columns -- with an FNV-1a 64 digest of the row bytes in that order.  Rows
// arrive in any order and any number; they are sorted in runs of bounded host
// memory, runs beyond the first go to temporary files next to the output, and
// the runs are merged into the file at the end.
//
// File: "GEMEMB01", u32 num_v, u32 0, u64 rows, u64 digest (32 bytes), then
// rows * num_v u32, little-endian.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gem_hip.h"

namespace gemh {

class EmbWriter {
 public:
  // run_bytes: host memory for one sorted run (0 = 256 MiB); max_rows: rows
  // beyond it make put() return false (0 = no bound).
  EmbWriter(std::string path, uint32_t num_v, uint64_t run_bytes, uint64_t max_rows);
  ~EmbWriter();
  EmbWriter(const EmbWriter &) = delete;
  EmbWriter &operator=(const EmbWriter &) = delete;

  // Append n rows (n * num_v ids).  False once more than max_rows were put or
  // a temporary file could not be written (status() says which).
  bool put(const uint32_t *rows, uint64_t n);
  // Sort and merge everything put so far into the output file.
  int finish(gem_export_result *out);
  int status() const { return status_; }
  // Removes the output and every temporary file (after a failed export).
  void discard();

 private:
  int flush_run();
  void sort_buffer();

  std::string path_;
  uint32_t nv_;
  uint64_t rows_per_run_, max_rows_, rows_ = 0;
  std::vector<uint32_t> buf_;
  std::vector<std::string> runs_;
  int status_ = GEM_OK;
  bool finished_ = false;
};

}  // namespace gemh
