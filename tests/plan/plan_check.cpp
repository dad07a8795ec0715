// Host-only check of the launch planners every entry point runs before its
// first kernel: the radix-sort chunk plan, the rebuild sub-tile plan and the
// row-scan piece plan.  Built with hipcc for the host alone (no device code
// is compiled or launched) by tests/test_plans.py; prints one line per
// failed invariant and "plans ok" when there is none.
#include <cstdio>
#include <vector>

#include "../../gpuquotientfilters_amd/csrc/gqf_kernels.hpp"
#include "../../gpuquotientfilters_amd/csrc/gqf_rebuild.hpp"

using namespace gqf;

static int fails = 0;
#define CHECK(c, ...)                           \
  do {                                          \
    if (!(c)) {                                 \
      ++fails;                                  \
      std::printf("FAIL %s: ", #c);             \
      std::printf(__VA_ARGS__);                 \
      std::printf("\n");                        \
    }                                           \
  } while (0)

int main() {
  std::vector<uint64_t> ns = {0, 1, 2, 63, 64, 4095, 4096, 4097, 8191, 8192, 8193, 65536, (1u << 20) + 1,
                              (1u << 22) - 1, 53687092, 60397976, 1u << 30, (1ull << 30) - 4096};
  for (uint32_t k = 0; k < 40; ++k) ns.push_back(((uint64_t)k * 2654435761u) % (1u << 27));
  for (uint64_t n : ns) {
    const SortPlan p = sort_plan((uint32_t)n);
    CHECK(p.G >= 1, "sort n=%llu G=%u", (unsigned long long)n, p.G);
    CHECK(p.chunk > 0 && p.chunk % SORT_TILE == 0, "sort n=%llu chunk=%u", (unsigned long long)n, p.chunk);
    CHECK((uint64_t)p.G * p.chunk >= n, "sort n=%llu does not cover", (unsigned long long)n);
    CHECK(n == 0 || (uint64_t)(p.G - 1) * p.chunk < n, "sort n=%llu idle chunk", (unsigned long long)n);
    CHECK(p.G <= 2048 || p.chunk == SORT_TILE, "sort n=%llu G=%u", (unsigned long long)n, p.G);
    CHECK(sort_workspace_bytes(n, 4, true) >= 4 * n * 4, "sort ws n=%llu", (unsigned long long)n);
    for (uint32_t per : {1u, 2u, 4u}) {
      const RbPlan r = rb_plan(n, per);
      CHECK(r.nsub == (n + RB_TILE - 1) / RB_TILE, "rb n=%llu nsub=%u", (unsigned long long)n, r.nsub);
      CHECK(r.G >= 1 && r.G <= RB_MAX_CHUNKS, "rb n=%llu G=%u", (unsigned long long)n, r.G);
      CHECK(r.chunk % RB_TILE == 0 && r.chunk >= per * RB_TILE, "rb n=%llu chunk=%u", (unsigned long long)n, r.chunk);
      CHECK((uint64_t)r.G * r.chunk >= n, "rb n=%llu does not cover", (unsigned long long)n);
    }
    const uint32_t rows = (uint32_t)(n / 64);
    const uint32_t pieces = (rows + SCAN_PIECE - 1) / SCAN_PIECE;
    CHECK(scan_rows_bytes(rows, 3) >= (size_t)pieces * 12, "scan rows=%u", rows);
  }
  if (!fails) std::printf("plans ok\n");
  return fails ? 1 : 0;
}
