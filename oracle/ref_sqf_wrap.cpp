// C entry points over the reference CPU_version SQF (quotientFilter.cpp) and
// the reference MT19937 (mt19937ar.cpp), compiled from the reference sources
// by oracle/Makefile into oracle/_ref/.  Test infrastructure only.
#include <stdint.h>
#include <stdlib.h>
#include "quotientFilter.h"  // reference header (quotientFilter.h:21-55)
#include "mt19937ar.h"       // reference header (mt19937ar.h:4-6)

extern "C" {
void* refsqf_new(unsigned q, unsigned r) {
  quotient_filter* f = new quotient_filter;
  initFilter(f, q, r);  // quotientFilter.cpp:34-40
  return f;
}
void refsqf_free(void* h) {
  quotient_filter* f = (quotient_filter*)h;
  free(f->table);
  delete f;
}
size_t refsqf_words(unsigned q, unsigned r) { return calcNumSlots(q, r); }
const uint32_t* refsqf_table(void* h) { return ((quotient_filter*)h)->table; }
uint32_t refsqf_insert(void* h, uint32_t key) { return insertItem((quotient_filter*)h, key); }
int32_t refsqf_may_contain(void* h, uint32_t key) { return mayContain((quotient_filter*)h, key); }
uint32_t refsqf_fnv(uint32_t key, uint32_t table_size) { return fnv_hash(key, table_size); }
void refsqf_insert_many(void* h, const uint32_t* keys, size_t n, uint32_t* out) {
  for (size_t i = 0; i < n; ++i) {
    uint32_t s = insertItem((quotient_filter*)h, keys[i]);
    if (out) out[i] = s;
  }
}
void refsqf_lookup_many(void* h, const uint32_t* keys, size_t n, int32_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = mayContain((quotient_filter*)h, keys[i]);
}
void refmt_fill(unsigned long seed, size_t n, uint32_t* out) {
  init_genrand(seed);
  for (size_t i = 0; i < n; ++i) out[i] = (uint32_t)genrand_int32();
}
}
