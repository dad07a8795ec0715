// A driver written against the reference host API (reference names, struct
// layouts and call sequence of the *Perf / *Test drivers), compiled against
// include/gqf_compat.hpp.  Device buffers come from the C ABI helpers so the
// driver builds with a plain C++ compiler.
//
//   compat_driver host                 sizes, struct layouts, hash values
//   compat_driver sqf  q r n seed      build/lookup/insert/delete checksums
//   compat_driver rsqf q n seed        (RBITS from the build) the same for the RSQF
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "gqf_compat.hpp"

static uint64_t fnv1a(const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

template <typename T>
static T* to_device(const std::vector<T>& v) {
  void* d = nullptr;
  if (gqf_malloc(&d, v.size() * sizeof(T) + 4) != GQF_OK) std::abort();
  if (!v.empty()) gqf_memcpy_htod(d, v.data(), v.size() * sizeof(T));
  return static_cast<T*>(d);
}
template <typename T>
static std::vector<T> to_host(const T* d, size_t n) {
  std::vector<T> v(n);
  if (n) gqf_memcpy_dtoh(v.data(), d, n * sizeof(T));
  return v;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  if (!std::strcmp(argv[1], "host")) {
    std::printf("sizeof_quotient_filter %zu %zu\n", sizeof(quotient_filter), sizeof(gqf_sqf));
    std::printf("offset_table %zu %zu\n", offsetof(quotient_filter, table), offsetof(gqf_sqf, table));
    std::printf("sizeof_cqf_gpu_block %zu\n", sizeof(cqf_gpu_block));
    const unsigned qs[] = {10, 18, 20, 26};
    for (unsigned q : qs) std::printf("slots %u 5 %zu\n", q, calcNumSlotsGPU(q, 5));
    for (unsigned q : qs) std::printf("slots %u 13 %zu\n", q, calcNumSlotsGPU(q, 13));
    for (unsigned q : qs) std::printf("blocks %u %zu\n", q, calcNumBlocksGPU(q));
    std::mt19937 g(1);
    for (int i = 0; i < 8; ++i) {
      const unsigned k = g();
      std::printf("ap %u %u %u\n", k, Normal_APHash(k, 1u << 31), Normal_APHash(k, 1u << 23));
    }
    return 0;
  }
  if (!std::strcmp(argv[1], "sqf") && argc == 6) {
    const unsigned q = (unsigned)std::atoi(argv[2]), r = (unsigned)std::atoi(argv[3]);
    const int n = std::atoi(argv[4]);
    std::mt19937 g((unsigned)std::atoi(argv[5]));
    std::vector<unsigned> keys(n + n / 4);
    for (auto& k : keys) k = g();
    quotient_filter qf;
    initFilterGPU(&qf, q, r);
    unsigned* d_keys = to_device(keys);
    const size_t tb = calcNumSlotsGPU(q, r);
    bulkBuildSegmentedLayouts(qf, n, d_keys, true);
    std::printf("build %016llx\n", (unsigned long long)fnv1a(to_host(qf.table, tb).data(), tb));
    std::vector<unsigned> out(keys.size());
    unsigned* d_out = to_device(out);
    launchUnsortedLookups(qf, (int)keys.size(), d_keys, d_out);
    auto o1 = to_host(d_out, keys.size());
    launchSortedLookups(qf, (int)keys.size(), d_keys, d_out);
    auto o2 = to_host(d_out, keys.size());
    std::printf("lookup %016llx %d\n", (unsigned long long)fnv1a(o1.data(), o1.size() * 4), o1 == o2 ? 1 : 0);
    insert(qf, n / 4, d_keys + n);
    std::printf("insert %016llx\n", (unsigned long long)fnv1a(to_host(qf.table, tb).data(), tb));
    superclusterDeletes(qf, n / 8, d_keys);
    std::printf("delete %016llx\n", (unsigned long long)fnv1a(to_host(qf.table, tb).data(), tb));
    gqf_free(d_keys);
    gqf_free(d_out);
    gqf_free(qf.table);
    return 0;
  }
  if (!std::strcmp(argv[1], "rsqf") && argc == 5) {
    const unsigned q = (unsigned)std::atoi(argv[2]);
    const int n = std::atoi(argv[3]);
    std::mt19937 g((unsigned)std::atoi(argv[4]));
    std::vector<unsigned> keys(n);
    for (auto& k : keys) k = g();
    countingQuotientFilterGPU cqf;
    initCQFGPU(&cqf, q);
    unsigned* d_keys = to_device(keys);
    std::vector<int> out(n);
    int* d_out = to_device(out);
    insertGPU(cqf, n, d_keys, d_out);
    std::vector<cqf_gpu_block> blocks(cqf.numBlocks);
    gqf_rsqf_dump(gqf_from(cqf), blocks.data());
    std::printf("insert %016llx\n", (unsigned long long)fnv1a(blocks.data(), blocks.size() * sizeof(cqf_gpu_block)));
    launchLookups(cqf, n, d_keys, d_out);
    auto o1 = to_host(d_out, n);
    launchUnsortedLookups(cqf, n, d_keys, d_out);
    auto o2 = to_host(d_out, n);
    std::printf("lookup %016llx %d\n", (unsigned long long)fnv1a(o1.data(), o1.size() * 4), o1 == o2 ? 1 : 0);
    gqf_free(d_keys);
    gqf_free(d_out);
    gqf_free(cqf.blocks);
    return 0;
  }
  return 2;
}
