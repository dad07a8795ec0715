// gqf_compat.hpp -- the reference host API, names and struct layouts
// unchanged, on top of the C ABI of libgqf.so (gqf.h).  Header-only C++.
//
// A driver written against the reference headers
//   standardQuotientFilter/GPU_version/quotientFilter.cuh:25-64
//   rankSelectQuotientFilter/GPU_version/RSQF.cuh:21-67
// compiles against this header instead and links -lgqf.  Every call blocks
// on the null stream and returns elapsed milliseconds, as in the reference.
//
// Error behaviour: the reference asserts in init (quotientFilter.cu:46-47)
// and aborts on library failures; here init errors and device errors print
// gqf_last_error() and abort().  A batch that does not fit the table (the
// reference writes out of bounds) prints the error, leaves the table
// unchanged and returns -1.0f.  mergeTwoFilters on filters of different
// shape prints and returns 0.0f (quotientFilter.cu:1749-1753).
//
// RSQF: RBITS (default 5, quotientFilter's RSQF.cuh:21) is a macro, define it
// before including this header to use another remainder width.  The device
// table behind countingQuotientFilterGPU::blocks is libgqf's own layout (see
// gqf_rsqf_table_bytes); callers only free it (hipFree), and
// printGPUFilter/gqf_rsqf_dump give the reference packed blocks.
#ifndef GQF_COMPAT_HPP
#define GQF_COMPAT_HPP

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gqf.h"

#ifndef NOT_FOUND
#define NOT_FOUND UINT_MAX
#endif
#ifndef RBITS
#define RBITS 5
#endif
#ifndef SLOTS_PER_BLOCK
#define SLOTS_PER_BLOCK 64
#endif

namespace gqf_compat_detail {
inline void die(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, gqf_last_error());
  std::abort();
}
// ms of a completed call; -1 after a capacity error (table unchanged)
inline float ms_or_fail(gqf_status s, float ms, const char* what) {
  if (s == GQF_OK) return ms;
  if (s == GQF_ERR_CAPACITY) {
    std::fprintf(stderr, "%s: %s\n", what, gqf_last_error());
    return -1.0f;
  }
  die(what);
  return 0.0f;
}
}  // namespace gqf_compat_detail

// ============================================================== standard QF
struct quotient_filter {
  unsigned int qbits;
  unsigned int rbits;
  unsigned int bytesPerElement;
  unsigned char* table;
};

inline gqf_sqf gqf_from(const quotient_filter& qf) {
  gqf_sqf g;
  g.qbits = qf.qbits;
  g.rbits = qf.rbits;
  g.bytesPerElement = qf.bytesPerElement;
  g.table = qf.table;
  return g;
}

// Table size in bytes (quotientFilter.cu:37-42).
inline size_t calcNumSlotsGPU(unsigned int q, unsigned int r) { return gqf_sqf_table_bytes(q, r); }

// AP hash (quotientFilter.cu:134-149): four byte rounds, most significant
// byte first, from state 0xAAAAAAAA; result modulo tableSize.
inline unsigned int Normal_APHash(unsigned int value, unsigned int tableSize) {
  unsigned int h = 0xAAAAAAAAu;
  for (int i = 0; i < 4; ++i) {
    const unsigned int b = (value >> (24 - 8 * i)) & 0xFFu;
    if ((i & 1) == 0) h ^= (h << 7) ^ b ^ (h >> 3);
    else h ^= ~((h << 11) ^ b ^ (h >> 5));
  }
  return h % tableSize;
}

// initFilterGPU (quotientFilter.cu:44-55); the table comes back zeroed.
inline void initFilterGPU(quotient_filter* qf, unsigned int q, unsigned int r) {
  gqf_sqf g;
  if (gqf_sqf_init(&g, q, r) != GQF_OK) gqf_compat_detail::die("initFilterGPU");
  qf->qbits = g.qbits;
  qf->rbits = g.rbits;
  qf->bytesPerElement = g.bytesPerElement;
  qf->table = g.table;
}

inline float launchSortedLookups(quotient_filter qf, int numValues, unsigned int* d_lookupValues,
                                 unsigned int* d_returnValues) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_lookup(gqf_from(qf), 1, numValues, d_lookupValues, d_returnValues, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "launchSortedLookups");
}

inline float launchUnsortedLookups(quotient_filter qf, int numValues, unsigned int* d_lookupValues,
                                   unsigned int* d_returnValues) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_lookup(gqf_from(qf), 0, numValues, d_lookupValues, d_returnValues, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "launchUnsortedLookups");
}

inline float insert(quotient_filter qf, int numValues, unsigned int* d_insertValues) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_insert(gqf_from(qf), numValues, d_insertValues, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "insert");
}

inline float bulkBuildParallelMerging(quotient_filter qf, int numValues, unsigned int* d_insertValues,
                                      bool NoDuplicates) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_bulk_build(gqf_from(qf), GQF_BUILD_PARALLEL_MERGING, numValues, d_insertValues,
                                          NoDuplicates ? 1 : 0, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "bulkBuildParallelMerging");
}

inline float bulkBuildSequentialShifts(quotient_filter qf, int numValues, unsigned int* d_insertValues,
                                       bool NoDuplicates) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_bulk_build(gqf_from(qf), GQF_BUILD_SEQUENTIAL_SHIFTS, numValues, d_insertValues,
                                          NoDuplicates ? 1 : 0, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "bulkBuildSequentialShifts");
}

// README.md names bulkBuildSequentialShifts "bulkBuildEarlyExit".
inline float bulkBuildEarlyExit(quotient_filter qf, int numValues, unsigned int* d_insertValues,
                                bool NoDuplicates) {
  return bulkBuildSequentialShifts(qf, numValues, d_insertValues, NoDuplicates);
}

inline float bulkBuildSegmentedLayouts(quotient_filter qf, int numValues, unsigned int* d_insertValues,
                                       bool NoDuplicates) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_bulk_build(gqf_from(qf), GQF_BUILD_SEGMENTED_LAYOUTS, numValues, d_insertValues,
                                          NoDuplicates ? 1 : 0, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "bulkBuildSegmentedLayouts");
}

inline float superclusterDeletes(quotient_filter qf, int numValues, unsigned int* d_deleteValues) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_delete(gqf_from(qf), numValues, d_deleteValues, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "superclusterDeletes");
}

inline float insertViaMerge(quotient_filter qf, unsigned int* d_insertedValues, int numOldValues,
                            unsigned int* d_newValues, int numNewValues, bool NoDuplicates) {
  float ms = 0.f;
  const gqf_status s = gqf_sqf_insert_via_merge(gqf_from(qf), d_insertedValues, numOldValues, d_newValues,
                                                numNewValues, NoDuplicates ? 1 : 0, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "insertViaMerge");
}

inline float mergeTwoFilters(quotient_filter qfilter1, quotient_filter qfilter2, bool NoDuplicates) {
  if (qfilter1.qbits != qfilter2.qbits || qfilter1.rbits != qfilter2.rbits) {
    std::printf("Error: two filters to be merged must have same number of quotient and remainder bits\n");
    return 0.0f;
  }
  float ms = 0.f;
  const gqf_status s = gqf_sqf_merge(gqf_from(qfilter1), gqf_from(qfilter2), NoDuplicates ? 1 : 0, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "mergeTwoFilters");
}

// printQuotientFilterGPU (quotientFilter.cu:383-405): rows of ten slots,
// metadata bits then remainders.
inline void printQuotientFilterGPU(quotient_filter* qf) {
  const size_t bytes = calcNumSlotsGPU(qf->qbits, qf->rbits);
  std::vector<unsigned char> h(bytes);
  if (gqf_memcpy_dtoh(h.data(), qf->table, bytes) != GQF_OK) gqf_compat_detail::die("printQuotientFilterGPU");
  const unsigned int bpe = qf->bytesPerElement;
  const unsigned int width = qf->rbits + 3;
  auto element = [&](size_t slot) {
    unsigned int e = 0;
    if (bpe == 0) {  // packed (r+3)-bit elements, MSB-first in 32-bit words
      for (unsigned int k = 0; k < width; ++k) {
        const size_t p = slot * width + k;
        unsigned int word;
        std::memcpy(&word, h.data() + 4 * (p / 32), 4);
        e = (e << 1) | ((word >> (31 - p % 32)) & 1u);
      }
      return e;
    }
    for (unsigned int b = 0; b < bpe; ++b) e = (e << 8) | h[slot * bpe + b];
    return e;
  };
  const int filterSize = (int)gqf_sqf_num_slots(qf->qbits);
  std::printf("Printing metadata and remainders:\n");
  for (int i = 0; i < filterSize / 10; i++) {
    for (int j = 0; j < 10; j++) std::printf("%u \t", element(10 * i + j) & 7u);
    std::printf("\n");
    for (int j = 0; j < 10; j++) std::printf("%u \t", element(10 * i + j) >> 3);
    std::printf("\n --------------------------------------------------------------------- \n");
  }
  std::printf("\n");
}

// ============================================================ rank-select QF
struct __attribute__((packed)) cqf_gpu_block {
  unsigned char offset;
  long unsigned int occupieds;
  long unsigned int runEnds;
  long unsigned int remainders[RBITS];
};

struct countingQuotientFilterGPU {
  unsigned int qbits;
  unsigned int numBlocks;
  cqf_gpu_block* blocks;
};

inline gqf_rsqf gqf_from(const countingQuotientFilterGPU& c) {
  gqf_rsqf g;
  g.qbits = c.qbits;
  g.numBlocks = c.numBlocks;
  g.blocks = c.blocks;
  g.rbits = RBITS;
  return g;
}

// calcNumBlocksGPU (RSQF.cu:27-30).
inline size_t calcNumBlocksGPU(unsigned int q) { return gqf_rsqf_num_blocks(q); }

inline unsigned int Normal_APHashGPU(unsigned int value, unsigned int maxHashValue) {
  return Normal_APHash(value, maxHashValue);
}

// initCQFGPU (RSQF.cu:32-40): allocated and zeroed.
inline void initCQFGPU(countingQuotientFilterGPU* cqf, unsigned int q) {
  gqf_rsqf g;
  if (gqf_rsqf_init(&g, q, RBITS) != GQF_OK) gqf_compat_detail::die("initCQFGPU");
  cqf->qbits = g.qbits;
  cqf->numBlocks = g.numBlocks;
  cqf->blocks = static_cast<cqf_gpu_block*>(g.blocks);
}

// launchLookups (RSQF.cu:411-456): the sorted path.
inline float launchLookups(countingQuotientFilterGPU cqf, int numValues, unsigned int* d_lookupValues,
                           int* slotValuesArray) {
  float ms = 0.f;
  const gqf_status s = gqf_rsqf_lookup(gqf_from(cqf), 1, numValues, d_lookupValues, slotValuesArray, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "launchLookups");
}

// launchUnsortedLookups (RSQF.cu:540-563), the RSQF overload.
inline float launchUnsortedLookups(countingQuotientFilterGPU cqf, int numValues, unsigned int* d_lookupValues,
                                   int* d_slotValuesArray) {
  float ms = 0.f;
  const gqf_status s = gqf_rsqf_lookup(gqf_from(cqf), 0, numValues, d_lookupValues, d_slotValuesArray, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "launchUnsortedLookups");
}

// insertGPU (RSQF.cu:882-988): d_returnValues receives each key's slot in
// the resulting (canonical) table, or -1 for a key that did not fit (the
// batch is then inserted in part, as the reference's insert leaves it).
inline float insertGPU(countingQuotientFilterGPU cqf, int numValues, unsigned int* d_insertValues,
                       int* d_returnValues) {
  float ms = 0.f;
  const gqf_status s =
      gqf_rsqf_insert_report(gqf_from(cqf), numValues, d_insertValues, d_returnValues, nullptr, nullptr, &ms);
  return gqf_compat_detail::ms_or_fail(s, ms, "insertGPU");
}

// printGPUFilter (RSQF.cu:127-172), from the reference packed layout.
inline void printGPUFilter(countingQuotientFilterGPU* cqf) {
  const gqf_rsqf g = gqf_from(*cqf);
  std::vector<cqf_gpu_block> h(cqf->numBlocks);
  if (gqf_rsqf_dump(g, h.data()) != GQF_OK) gqf_compat_detail::die("printGPUFilter");
  auto rem = [&](int blk, int slot) {
    const unsigned int bit = (unsigned int)slot * RBITS;
    unsigned long long v = 0;
    for (unsigned int b = 0; b < RBITS; ++b) {
      const unsigned int at = bit + b;
      v |= ((h[blk].remainders[at / 64] >> (at % 64)) & 1ull) << b;
    }
    return (unsigned int)v;
  };
  std::printf("Filter contents:\n");
  for (int i = 0; i < (int)cqf->numBlocks; i++) {
    std::printf("block: %i\t", i);
    std::printf("offset: %u\n", h[i].offset);
    const int rowsPerBlock = SLOTS_PER_BLOCK / 10;
    auto row = [&](int from, int count) {
      std::printf("\noccupieds:\t");
      for (int k = 0; k < count; k++) std::printf("%u\t", (unsigned int)((h[i].occupieds >> (from + k)) & 1ul));
      std::printf("\nrunEnds:\t");
      for (int k = 0; k < count; k++) std::printf("%u\t", (unsigned int)((h[i].runEnds >> (from + k)) & 1ul));
      std::printf("\nremainders:\t");
      for (int k = 0; k < count; k++) std::printf("%u\t", rem(i, from + k));
      std::printf("\n");
    };
    for (int j = 0; j < rowsPerBlock; j++) row(10 * j, 10);
    if (SLOTS_PER_BLOCK % 10 != 0) row(rowsPerBlock * 10, SLOTS_PER_BLOCK % 10);
    std::printf("\n --------------------------------------------------------------------- \n");
  }
}

#endif  // GQF_COMPAT_HPP
