"""The C++ examples of INTEGRATION.md compile and link against
include/gem_hip.h and the built library (no GPU call: nothing is run).  Each
example is placed in its own scope of a function whose parameters are the
values the surrounding text assumes (a context, a query, options, result
arrays, the reference loader's host graphs)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gem_amd", "_lib")

HARNESS = r"""
#include <cstdio>
#include <map>
#include <utility>
#include <vector>
#include "gem_hip.h"

// the fields of the reference loader's host graph the examples read
struct RefHostGraph {
  uint32_t num_v = 0, num_e = 0;
  std::vector<uint32_t> offsets_, neighbors_, vLabels_;
  std::map<uint32_t, std::pair<uint32_t, uint32_t>> evv;
};
static const char *bits32(const uint64_t *) { return ""; }

void example(gem_ctx *ctx, gem_query_view q, uint32_t T, uint64_t cap, std::vector<gem_query_view> queries,
             uint32_t nq, gem_batch_opts o, std::vector<gem_join_result> res, gem_csr_view v, gem_graph *g,
             const float *trained_params, float margin, uint32_t *chosen, RefHostGraph hq, RefHostGraph hg) {
  (void)ctx; (void)q; (void)T; (void)cap; (void)queries; (void)nq; (void)o; (void)res; (void)v; (void)g;
  (void)trained_params; (void)margin; (void)chosen; (void)hq; (void)hg;
  {
%s
  }
}

int main() { return 0; }
"""


def _examples():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_integration_examples_compile_and_link(tmp_path):
    if not os.path.exists(os.path.join(LIB, "libgem_hip.so")):
        pytest.skip("library not built")
    blocks = _examples()
    assert len(blocks) >= 6
    for i, b in enumerate(blocks):
        body = "\n".join(line for line in b.splitlines() if not line.startswith("#include"))
        src = tmp_path / f"example{i}.cpp"
        src.write_text(HARNESS % body)
        exe = tmp_path / f"example{i}"
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
                            "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIB, "-lgem_hip",
                            f"-Wl,-rpath,{LIB}", "-Wl,-rpath,/opt/rocm/lib"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, f"example {i}:\n{b}\n{r.stderr[-3000:]}"
