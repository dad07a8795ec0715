// C entry points over the reference CPU_version RSQF (RSQF.cpp, RBITS as the
// reference header defines it), compiled from the reference sources by
// oracle/Makefile into oracle/_ref/.  Test infrastructure only.
#include <stdint.h>
#include <stdlib.h>
#include "RSQF.h"  // reference header (RSQF.h:20-59)

unsigned int rankBit(long long unsigned int bitArray, unsigned int index);  // RSQF.cpp:185
unsigned int selectBit(long long unsigned int x, unsigned int k);           // RSQF.cpp:290

extern "C" {
unsigned refrsqf_rbits(void) { return RBITS; }
size_t refrsqf_block_bytes(void) { return sizeof(cqf_block); }
void* refrsqf_new(unsigned q) {
  countingQuotientFilter* f = new countingQuotientFilter;
  initCQF(f, q);  // RSQF.cpp:27-33
  return f;
}
void refrsqf_free(void* h) {
  countingQuotientFilter* f = (countingQuotientFilter*)h;
  free(f->blocks);
  delete f;
}
size_t refrsqf_num_blocks(unsigned q) { return calcNumBlocks(q); }
const uint8_t* refrsqf_blocks(void* h) { return (const uint8_t*)((countingQuotientFilter*)h)->blocks; }
int32_t refrsqf_insert(void* h, uint32_t key) { return insert((countingQuotientFilter*)h, key); }
int32_t refrsqf_lookup(void* h, uint32_t key) { return lookup((countingQuotientFilter*)h, key); }
uint32_t refrsqf_ap(uint32_t key, uint32_t max_hash) { return Normal_APHash(key, max_hash); }
uint32_t refrsqf_rank(uint64_t x, unsigned i) { return rankBit(x, i); }
uint32_t refrsqf_select(uint64_t x, unsigned k) { return selectBit(x, k); }
void refrsqf_insert_many(void* h, const uint32_t* keys, size_t n, int32_t* out) {
  for (size_t i = 0; i < n; ++i) {
    int32_t s = insert((countingQuotientFilter*)h, keys[i]);
    if (out) out[i] = s;
  }
}
void refrsqf_lookup_many(void* h, const uint32_t* keys, size_t n, int32_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = lookup((countingQuotientFilter*)h, keys[i]);
}
}
