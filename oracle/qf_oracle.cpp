// qf_oracle.cpp -- CPU oracle for the quotient-filter engine.
//
// TEST INFRASTRUCTURE ONLY: the product library never links or calls this.
// Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
//
// Every function restates the reference algorithm it cites.  Reference paths:
//   quotientFilter.cu  = standardQuotientFilter/GPU_version/quotientFilter.cu
//   quotientFilter.cpp = standardQuotientFilter/CPU_version/quotientFilter.cpp
//   RSQF.cpp / RSQF.cu = rankSelectQuotientFilter/{CPU,GPU}_version/
// Deviations from the reference are bounds checks only (reads past the last
// slot/block read as empty; writes past it become ORC_ERR_CAPACITY).

#include "qf_oracle.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

inline uint32_t low_mask32(unsigned n) { return n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u); }
inline uint64_t low_mask64(unsigned n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

// ------------------------------------------------------------------ SQF
// Slot codec of the GPU_version: (r+3)/8 bytes per slot, big-endian
// (quotientFilter.cu:151-170).  element = rem<<3 | occupied<<2 | continuation<<1 | shifted.
// For widths the GPU_version does not support ((r+3) % 8 != 0) the model
// keeps the same elements in ceil((r+3)/8) bytes (a "wide" table); the
// packed (r+3)-bit layout is compared against it element by element.
struct Sqf {
  uint8_t* t;
  unsigned q, r, bpe;
  uint32_t S;  // number of slots, floor(2^q * 1.1) (quotientFilter.cu:174)
  uint32_t get(uint64_t i) const {
    if (i >= S) return 0;  // bounds check: past-the-end reads as empty
    const uint8_t* p = t + i * bpe;
    uint32_t e = p[0];
    for (unsigned k = 1; k < bpe; ++k) e = (e << 8) | p[k];
    return e;
  }
  void set(uint64_t i, uint32_t v) {
    uint8_t* p = t + i * bpe;
    for (unsigned k = 0; k < bpe; ++k) p[k] = (uint8_t)(v >> (8 * (bpe - 1 - k)));
  }
};
inline bool occ(uint32_t e) { return e & 4; }
inline bool cont(uint32_t e) { return e & 2; }
inline bool shf(uint32_t e) { return e & 1; }
inline bool empty_(uint32_t e) { return (e & 7) == 0; }

Sqf mk_sqf(uint8_t* t, unsigned q, unsigned r) {
  Sqf f;
  f.t = t;
  f.q = q;
  f.r = r;
  f.bpe = (r + 3 + 7) / 8;
  f.S = orc_sqf_num_slots(q);
  return f;
}

// findRunStartGPU (quotientFilter.cu:172-196) == findRunStart (quotientFilter.cpp:176-199)
uint32_t find_run_start(const Sqf& f, uint32_t fq) {
  uint32_t b = fq;
  while (b > 0 && shf(f.get(b))) b--;
  uint32_t s = b;
  while (b != fq) {
    do { s++; } while (cont(f.get(s)) && s < f.S);
    do { b++; } while (!occ(f.get(b)) && b < f.S);
  }
  return s;
}

// ------------------------------------------------------------------ RSQF
// Reference packed block (RSQF.cuh:24-30 / RSQF.h:26-32) with runtime r:
//   u8 offset; u64 occupieds; u64 runEnds; u64 remainders[r];  (packed, LE)
struct Rsqf {
  uint8_t* t;
  unsigned q, r;
  uint32_t NB;
  size_t BB;
  uint8_t offset(uint32_t b) const { return t[b * BB]; }
  void set_offset(uint32_t b, uint8_t v) { t[b * BB] = v; }
  uint64_t occ(uint32_t b) const { uint64_t v; memcpy(&v, t + b * BB + 1, 8); return v; }
  uint64_t ends(uint32_t b) const { uint64_t v; memcpy(&v, t + b * BB + 9, 8); return v; }
  void set_occ(uint32_t b, uint64_t v) { memcpy(t + b * BB + 1, &v, 8); }
  void set_ends(uint32_t b, uint64_t v) { memcpy(t + b * BB + 9, &v, 8); }
  // getRemainder / setRemainder (RSQF.cpp:89-123): bits [s*r, s*r+r) of the
  // little-endian remainder words, which is the same as LSB-first bit order
  // over the packed remainder bytes.
  uint32_t rem(uint32_t b, uint32_t s) const {
    const uint8_t* base = t + b * BB + 17;
    uint64_t bit = (uint64_t)s * r;
    uint32_t v = 0;
    for (unsigned k = 0; k < r; ++k) {
      uint64_t p = bit + k;
      v |= (uint32_t)((base[p >> 3] >> (p & 7)) & 1u) << k;
    }
    return v;
  }
  void set_rem(uint32_t b, uint32_t s, uint32_t val) {
    uint8_t* base = t + b * BB + 17;
    uint64_t bit = (uint64_t)s * r;
    for (unsigned k = 0; k < r; ++k) {
      uint64_t p = bit + k;
      uint8_t m = (uint8_t)(1u << (p & 7));
      if ((val >> k) & 1u) base[p >> 3] |= m; else base[p >> 3] &= (uint8_t)~m;
    }
  }
};

Rsqf mk_rsqf(uint8_t* t, unsigned q, unsigned r) {
  Rsqf f;
  f.t = t;
  f.q = q;
  f.r = r;
  f.NB = orc_rsqf_num_blocks(q);
  f.BB = orc_rsqf_block_bytes(r);
  return f;
}

inline bool bit_of(uint64_t x, uint32_t s) { return s < 64 && ((x >> s) & 1ull); }

}  // namespace

extern "C" {

// ================================================================== keys
// The reference generator (mt19937ar.cpp init_genrand/genrand_int32) is the
// standard MT19937; std::mt19937 produces the identical stream.
void orc_mt19937(uint32_t seed, uint64_t skip, size_t n, uint32_t* out) {
  std::mt19937 g(seed);
  g.discard(skip);
  for (size_t i = 0; i < n; ++i) out[i] = (uint32_t)g();
}

// ================================================================== hashes
// Normal_APHash rounds (quotientFilter.cu:134-149, RSQF.cu:174-189,
// RSQF.cpp:168-183 -- identical bodies) from an arbitrary initial state.
uint32_t orc_ap_rounds(uint32_t key, uint32_t state) {
  uint8_t p[4] = {(uint8_t)(key >> 24), (uint8_t)(key >> 16), (uint8_t)(key >> 8), (uint8_t)key};
  uint32_t h = state;
  for (int i = 0; i < 4; ++i)
    h ^= ((i & 1) == 0) ? ((h << 7) ^ p[i] ^ (h >> 3)) : (~((h << 11) ^ p[i] ^ (h >> 5)));
  return h;
}

// Fingerprint of width q+r.  q+r <= 31: AP(key) mod 2^(q+r) exactly as the
// reference.  q+r == 32: AP(key).  q+r > 32 (undefined in the reference, whose
// modulus 1<<(q+r) overflows): the MurmurHash3 64-bit finalizer (fmix64,
// Appleby, public domain) of (AP(key) << 32 | key), truncated to q+r bits.
// The input word is injective in the key, so the result is uniform; AP alone
// (or any function of it) reaches only about a third of the 32-bit range.
uint64_t orc_fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
uint64_t orc_fp_ap(uint32_t key, unsigned q, unsigned r) {
  unsigned bits = q + r;
  uint32_t lo = orc_ap_rounds(key, 0xAAAAAAAAu);
  if (bits <= 32) return lo & low_mask32(bits);
  return orc_fmix64(((uint64_t)lo << 32) | key) & low_mask64(bits);
}

// fnv_hash (quotientFilter.cpp:97-112): FNV-1, multiply then xor, big-endian bytes.
uint64_t orc_fp_fnv(uint32_t key, unsigned q, unsigned r) {
  uint8_t p[4] = {(uint8_t)(key >> 24), (uint8_t)(key >> 16), (uint8_t)(key >> 8), (uint8_t)key};
  uint32_t h = 2166136261u;
  for (int i = 0; i < 4; ++i) h = (h * 16777619u) ^ p[i];
  return h & low_mask32(q + r);
}

void orc_fingerprints(const uint32_t* keys, size_t n, unsigned q, unsigned r, int hash_kind,
                      uint64_t* out) {
  for (size_t i = 0; i < n; ++i)
    out[i] = hash_kind == 1 ? orc_fp_fnv(keys[i], q, r) : orc_fp_ap(keys[i], q, r);
}

// ================================================================== SQF geometry
// numElements = (1 << q) * 1.1 (quotientFilter.cu:174,535)
uint32_t orc_sqf_num_slots(unsigned q) { return (uint32_t)((double)(1u << q) * 1.1); }

// calcNumSlotsGPU (quotientFilter.cu:37-42): a byte count despite the name.
size_t orc_sqf_table_bytes(unsigned q, unsigned r) {
  // widths the GPU_version does not support: S slots of ceil((r+3)/8) bytes
  if ((r + 3) % 8 != 0) return (size_t)orc_sqf_num_slots(q) * ((r + 3 + 7) / 8);
  size_t tableBits = (size_t)(1u << q) * (r + 3);
  size_t tableSlots = tableBits / 8;
  return (size_t)((double)tableSlots * 1.1);
}

// calcNumSlots (quotientFilter.cpp:26-32): uint32 words of the CPU_version table.
size_t orc_sqf_cpu_words(unsigned q, unsigned r) {
  size_t tableBits = (size_t)(1u << q) * (r + 3);
  size_t tableSlots = tableBits / 32;
  if ((tableSlots % 32) > 1) tableSlots += 1;
  return (size_t)((double)tableSlots * 1.1);
}

uint32_t orc_sqf_get(const uint8_t* table, unsigned r, uint32_t slot) {
  unsigned bpe = (r + 3 + 7) / 8;
  uint32_t e = table[(size_t)slot * bpe];
  for (unsigned k = 1; k < bpe; ++k) e = (e << 8) | table[(size_t)slot * bpe + k];
  return e;
}

void orc_sqf_set(uint8_t* table, unsigned r, uint32_t slot, uint32_t elem) {
  unsigned bpe = (r + 3 + 7) / 8;
  for (unsigned k = 0; k < bpe; ++k)
    table[(size_t)slot * bpe + k] = (uint8_t)(elem >> (8 * (bpe - 1 - k)));
}

void orc_sqf_decode_bytes(const uint8_t* table, unsigned r, uint32_t nslots, uint32_t* elems) {
  for (uint32_t i = 0; i < nslots; ++i) elems[i] = orc_sqf_get(table, r, i);
}

// getElement of the CPU_version (quotientFilter.cpp:114-130): (r+3)-bit slots
// packed MSB-first into uint32 words.  The CPU table can hold fewer than S
// slots (calcNumSlots rounds differently); bits past nwords read as 0.
void orc_sqf_decode_words(const uint32_t* words, size_t nwords, unsigned r, uint32_t nslots,
                          uint32_t* elems) {
  unsigned w = r + 3;
  for (uint32_t i = 0; i < nslots; ++i) {
    uint64_t bit = (uint64_t)i * w;
    uint32_t v = 0;
    for (unsigned k = 0; k < w; ++k) {
      uint64_t p = bit + k;
      uint32_t word = p / 32 < nwords ? words[p / 32] : 0u;
      v = (v << 1) | ((word >> (31 - (p % 32))) & 1u);
    }
    elems[i] = v;
  }
}

// ================================================================== SQF canonical layout
// Closed form of every SQF end state: runs in quotient order, remainders sorted
// within a run, entry i at loc_i = max(fq_i, loc_{i-1}+1).  Matches the write
// rules of setMetadata (quotientFilter.cu:700-718), segmentedQFWrite
// (quotientFilter.cu:1430-1443) and setOccupiedBits (quotientFilter.cu:728-735).
int orc_sqf_canon(const uint64_t* fps, size_t n, unsigned q, unsigned r, int no_dup,
                  uint8_t* table) {
  std::vector<uint64_t> v(fps, fps + n);
  std::sort(v.begin(), v.end());
  if (no_dup) v.erase(std::unique(v.begin(), v.end()), v.end());
  Sqf f = mk_sqf(table, q, r);
  size_t bytes = orc_sqf_table_bytes(q, r);
  std::vector<uint32_t> loc(v.size());
  uint64_t prev = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    uint64_t fq = (v[i] >> r) & low_mask64(q);
    uint64_t l = (i == 0) ? fq : std::max<uint64_t>(fq, prev + 1);
    if (l >= f.S) return ORC_ERR_CAPACITY;
    loc[i] = (uint32_t)l;
    prev = l;
  }
  memset(table, 0, bytes);
  for (size_t i = 0; i < v.size(); ++i) {
    uint32_t fq = (uint32_t)((v[i] >> r) & low_mask64(q));
    uint32_t fr = (uint32_t)(v[i] & low_mask64(r));
    uint32_t e = fr << 3;
    if (i > 0 && ((v[i - 1] >> r) & low_mask64(q)) == fq) e |= 2;
    if (loc[i] != fq) e |= 1;
    f.set(loc[i], e);
  }
  for (size_t i = 0; i < v.size(); ++i) {
    uint32_t fq = (uint32_t)((v[i] >> r) & low_mask64(q));
    f.set(fq, f.get(fq) | 4);
  }
  return ORC_OK;
}

// ================================================================== SQF faithful ops
// lookUp / hashAndLookUp (quotientFilter.cu:222-259, 320-357) == mayContain
// (quotientFilter.cpp:202-234): slot of the FIRST matching remainder of the run.
uint32_t orc_sqf_lookup_fp(const uint8_t* table, unsigned q, unsigned r, uint64_t fp) {
  Sqf f = mk_sqf(const_cast<uint8_t*>(table), q, r);
  uint32_t fq = (uint32_t)((fp >> r) & low_mask64(q));
  uint32_t fr = (uint32_t)(fp & low_mask64(r));
  if (!occ(f.get(fq))) return ORC_NOT_FOUND;
  uint32_t s = find_run_start(f, fq);
  do {
    uint32_t rem = f.get(s) >> 3;
    if (rem == fr) return s;
    if (rem > fr) return ORC_NOT_FOUND;
    s++;
  } while (cont(f.get(s)));
  return ORC_NOT_FOUND;
}

// insertItemGPU (quotientFilter.cu:449-531) + insertItemHereGPU (:198-220)
// == insertItem/insertItemHere (quotientFilter.cpp:236-325).  Deduplicating:
// a fingerprint already present is not inserted again.
uint32_t orc_sqf_insert_fp(uint8_t* table, unsigned q, unsigned r, uint64_t fp, int* status) {
  Sqf f = mk_sqf(table, q, r);
  if (status) *status = ORC_OK;
  uint32_t fq = (uint32_t)((fp >> r) & low_mask64(q));
  uint32_t fr = (uint32_t)(fp & low_mask64(r));
  uint32_t canon = f.get(fq);
  uint32_t ne = fr << 3;
  if (empty_(canon)) {
    f.set(fq, ne | 4);
    return fq;
  }
  // bounds check before any write: the shift ends at the first empty slot at
  // or after fq (clusters hold no empty slots), which must exist below S.
  {
    uint32_t e = fq;
    while (e < f.S && !empty_(f.get(e))) e++;
    if (e >= f.S) {
      // a duplicate would not shift anything; look it up first
      if (occ(canon)) {
        uint32_t hit = orc_sqf_lookup_fp(table, q, r, fp);
        if (hit != ORC_NOT_FOUND) return hit;
      }
      if (status) *status = ORC_ERR_CAPACITY;
      return ORC_NOT_FOUND;
    }
  }
  if (!occ(canon)) f.set(fq, canon | 4);
  uint32_t runStart = find_run_start(f, fq);
  uint32_t s = runStart;
  if (occ(canon)) {
    do {
      uint32_t rem = f.get(s) >> 3;
      if (rem == fr) return s;
      if (rem > fr) break;
      s++;
    } while (cont(f.get(s)));
    if (s == runStart) f.set(runStart, f.get(runStart) | 2);
    else ne |= 2;
  }
  if (s != fq) ne |= 1;
  // insertItemHereGPU
  uint32_t idx = s, cur = ne;
  bool done = false;
  while (!done) {
    uint32_t prev = f.get(idx);
    done = empty_(prev);
    prev |= 1;
    if (occ(prev)) {
      prev &= ~4u;
      cur |= 4;
    }
    f.set(idx, cur);
    cur = prev;
    idx++;
  }
  return s;
}

// deleteItemGPU (quotientFilter.cu:1076-1200), run sequentially.  Removes one
// copy of fp; reads past the last slot read as empty.
int orc_sqf_delete_fp(uint8_t* table, unsigned q, unsigned r, uint64_t fp) {
  Sqf f = mk_sqf(table, q, r);
  uint32_t fq = (uint32_t)((fp >> r) & low_mask64(q));
  uint32_t fr = (uint32_t)(fp & low_mask64(r));
  if (!occ(f.get(fq))) return 0;
  uint32_t b = fq;
  while (b > 0 && shf(f.get(b))) b--;
  uint32_t s = b;
  while (b != fq) {
    do { s++; } while (cont(f.get(s)));
    do { b++; } while (!occ(f.get(b)) && b < f.S);
  }
  uint32_t runStart = s;
  uint32_t rem;
  do {
    rem = f.get(s) >> 3;
    if (rem == fr) break;
    if (rem > fr) return 0;
    s++;
  } while (cont(f.get(s)));
  if (rem != fr) return 0;
  if (!cont(f.get(s + 1))) {
    do { b++; } while (!occ(f.get(b)) && b < f.S);
  }
  if (s == runStart) {
    if (!cont(f.get(s + 1))) f.set(fq, f.get(fq) & ~4u);
    else f.set(s + 1, f.get(s + 1) & ~2u);
  }
  while (shf(f.get(s + 1))) {
    if (b == s) f.set(s + 1, f.get(s + 1) & ~1u);
    do {
      uint32_t next = f.get(s + 1);
      if (occ(f.get(s))) f.set(s, next | 4);
      else f.set(s, next & ~4u);
      s++;
    } while (cont(f.get(s + 1)));
    do { b++; } while (!occ(f.get(b)) && b < f.S);
  }
  f.set(s, 0);
  return 1;
}

// extractQuotients (quotientFilter.cu:1561-1600) over the S slots (the
// reference passes the byte count instead, which over-reads for r = 13).
size_t orc_sqf_extract(const uint8_t* table, unsigned q, unsigned r, uint64_t* out) {
  Sqf f = mk_sqf(const_cast<uint8_t*>(table), q, r);
  size_t k = 0;
  for (uint32_t idx = 0; idx < f.S; ++idx) {
    uint32_t e = f.get(idx);
    if (empty_(e)) continue;
    if (!shf(e)) {
      out[k++] = ((uint64_t)idx << r) | (e >> 3);
      continue;
    }
    uint32_t b = idx;
    do { b--; } while (b > 0 && shf(f.get(b)));
    uint32_t s = b;
    while (s <= idx) {
      do { s++; } while (cont(f.get(s)));
      if (s > idx) break;
      do { b++; } while (!occ(f.get(b)) && b < f.S);
    }
    out[k++] = ((uint64_t)b << r) | (e >> 3);
  }
  return k;
}

static uint64_t fp_of(uint32_t key, unsigned q, unsigned r, int hash_kind) {
  return hash_kind == 1 ? orc_fp_fnv(key, q, r) : orc_fp_ap(key, q, r);
}

void orc_sqf_lookup_keys(const uint8_t* table, unsigned q, unsigned r, int hash_kind,
                         const uint32_t* keys, size_t n, uint32_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = orc_sqf_lookup_fp(table, q, r, fp_of(keys[i], q, r, hash_kind));
}

int orc_sqf_insert_keys(uint8_t* table, unsigned q, unsigned r, int hash_kind,
                        const uint32_t* keys, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    int st;
    orc_sqf_insert_fp(table, q, r, fp_of(keys[i], q, r, hash_kind), &st);
    if (st != ORC_OK) return st;
  }
  return ORC_OK;
}

size_t orc_sqf_delete_keys(uint8_t* table, unsigned q, unsigned r, int hash_kind,
                           const uint32_t* keys, size_t n) {
  size_t k = 0;
  for (size_t i = 0; i < n; ++i) k += orc_sqf_delete_fp(table, q, r, fp_of(keys[i], q, r, hash_kind));
  return k;
}

// ================================================================== RSQF
// calcNumBlocksGPU (RSQF.cu:27-30)
uint32_t orc_rsqf_num_blocks(unsigned q) { return (uint32_t)((((uint64_t)1 << q) + 63) / 64 + 1); }
size_t orc_rsqf_block_bytes(unsigned r) { return 1 + 8 + 8 + 8 * (size_t)r; }
size_t orc_rsqf_table_bytes(unsigned q, unsigned r) {
  return (size_t)orc_rsqf_num_blocks(q) * orc_rsqf_block_bytes(r);
}

// rankBit (RSQF.cpp:185-189): popcount of bits [0, i].
uint32_t orc_rank(uint64_t x, unsigned i) { return (uint32_t)__builtin_popcountll(x & low_mask64(i + 1)); }

// selectBit (RSQF.cpp:290-312), stated by its definition: k == 0 -> 0;
// position of the k-th set bit (1-based); UINT_MAX if x has fewer than k.
uint32_t orc_select(uint64_t x, unsigned k) {
  if (k == 0) return 0;
  unsigned seen = 0;
  for (unsigned i = 0; i < 64; ++i)
    if ((x >> i) & 1ull)
      if (++seen == k) return i;
  return ORC_NOT_FOUND;
}

uint32_t orc_rsqf_get_rem(const uint8_t* table, unsigned r, uint32_t block, unsigned slot) {
  Rsqf f;
  f.t = const_cast<uint8_t*>(table);
  f.r = r;
  f.BB = orc_rsqf_block_bytes(r);
  return f.rem(block, slot);
}

// Canonical RSQF layout: rem[loc_i] = fr_i; occupieds bit fq_i; runEnds bit at
// the last entry of each run; offset_j = (E_j - 64 j + 1) mod 256 when the
// largest location E_j of an entry with fq < 64 j lies in block j or later,
// else 0.  Equal to the reference insert (RSQF.cpp:411-569) fed the
// fingerprints in ascending order.
int orc_rsqf_canon(const uint64_t* fps, size_t n, unsigned q, unsigned r, uint8_t* table) {
  std::vector<uint64_t> v(fps, fps + n);
  std::sort(v.begin(), v.end());
  Rsqf f = mk_rsqf(table, q, r);
  uint64_t cap = (uint64_t)f.NB * 64;
  std::vector<uint64_t> loc(v.size()), fqs(v.size());
  uint64_t prev = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    uint64_t fq = (v[i] >> r) & low_mask64(q);
    uint64_t l = i == 0 ? fq : std::max<uint64_t>(fq, prev + 1);
    if (l >= cap) return ORC_ERR_CAPACITY;
    loc[i] = l;
    fqs[i] = fq;
    prev = l;
  }
  memset(table, 0, orc_rsqf_table_bytes(q, r));
  for (size_t i = 0; i < v.size(); ++i) {
    uint32_t b = (uint32_t)(loc[i] / 64), s = (uint32_t)(loc[i] % 64);
    f.set_rem(b, s, (uint32_t)(v[i] & low_mask64(r)));
    f.set_occ((uint32_t)(fqs[i] / 64), f.occ((uint32_t)(fqs[i] / 64)) | (1ull << (fqs[i] % 64)));
    if (i + 1 == v.size() || fqs[i + 1] != fqs[i]) f.set_ends(b, f.ends(b) | (1ull << s));
  }
  // offsets: walk blocks with a pointer to the last entry whose fq < 64 j
  size_t p = 0;
  for (uint32_t j = 0; j < f.NB; ++j) {
    uint64_t lim = (uint64_t)j * 64;
    while (p < v.size() && fqs[p] < lim) p++;
    if (p == 0) continue;
    uint64_t E = loc[p - 1];
    if (E >= lim) f.set_offset(j, (uint8_t)((E - lim + 1) & 0xFF));
  }
  return ORC_OK;
}

// lookup (RSQF.cpp:314-381) == lookupGPU / hashAndLookupGPU (RSQF.cu:320-399,
// 459-538): the LAST matching slot of the run (scan backwards from the run
// end), else NOT_FOUND.  Faithful on any table, including ones with holes.
// Bounds: the reference reads block numBlocks (its check is `>`); here any
// block >= numBlocks ends the search with NOT_FOUND.
uint32_t orc_rsqf_lookup_fp(const uint8_t* table, unsigned q, unsigned r, uint64_t fp) {
  Rsqf f = mk_rsqf(const_cast<uint8_t*>(table), q, r);
  uint32_t fq = (uint32_t)((fp >> r) & low_mask64(q));
  uint32_t fr = (uint32_t)(fp & low_mask64(r));
  uint32_t blockNum = fq / 64, slotNum = fq % 64;
  uint64_t occs = f.occ(blockNum);
  if (!bit_of(occs, slotNum)) return ORC_NOT_FOUND;
  uint8_t blockOffset = f.offset(blockNum);
  uint32_t rank = orc_rank(occs, slotNum);
  uint32_t endOfRun = orc_select(f.ends(blockNum) & ~low_mask64(blockOffset), rank);
  while (endOfRun == ORC_NOT_FOUND) {
    rank -= (uint32_t)__builtin_popcountll(f.ends(blockNum) & ~low_mask64(blockOffset));
    if ((int)blockOffset - 64 > 0) blockOffset = (uint8_t)(blockOffset - 64);
    else blockOffset = 0;
    blockNum++;
    if (blockNum >= f.NB) return ORC_NOT_FOUND;
    endOfRun = orc_select(f.ends(blockNum) & ~low_mask64(blockOffset), rank);
  }
  uint32_t cur = f.rem(blockNum, endOfRun);
  uint32_t slot = endOfRun;
  do {
    if (cur == fr) return blockNum * 64 + slot;
    if (cur < fr) return ORC_NOT_FOUND;
    if (slot > 0) {
      slot--;
    } else {
      slot = 63;
      if (blockNum == 0) return ORC_NOT_FOUND;
      blockNum--;
    }
    cur = f.rem(blockNum, slot);
  } while (!bit_of(f.ends(blockNum), slot) && (blockNum * 64 + slot) >= fq);
  return ORC_NOT_FOUND;
}

// findFirstUnusedSlot (RSQF.cpp:383-409), including the rank==0 -> select =
// offset quirk that can skip a free slot.
static uint32_t rsqf_first_unused(Rsqf& f, int* blockNum, uint32_t startSlot) {
  uint32_t cur = startSlot;
  uint64_t occs = f.occ(*blockNum), ends = f.ends(*blockNum);
  uint8_t offset = f.offset(*blockNum);
  uint32_t rank = orc_rank(occs, startSlot);
  uint32_t sel = orc_select(ends & ~low_mask64(offset), rank);
  if (rank == 0) sel = offset;
  while (cur <= sel) {
    if (sel == ORC_NOT_FOUND || sel == 63) {
      (*blockNum)++;
      if ((uint32_t)*blockNum >= f.NB) return ORC_NOT_FOUND;
      occs = f.occ(*blockNum);
      ends = f.ends(*blockNum);
      offset = f.offset(*blockNum);
      sel = (uint32_t)((int)offset - 1);
    }
    cur = sel + 1;
    rank = orc_rank(occs, cur);
    sel = orc_select(ends & ~low_mask64(offset), rank);
  }
  return cur;
}

// insert (RSQF.cpp:411-569), sequential and faithful (duplicates are stored).
// Returns the slot written, or NOT_FOUND when the reference would fail or
// would index past its arrays.
uint32_t orc_rsqf_insert_fp(uint8_t* table, unsigned q, unsigned r, uint64_t fp) {
  Rsqf f = mk_rsqf(table, q, r);
  uint32_t fq = (uint32_t)((fp >> r) & low_mask64(q));
  uint32_t fr = (uint32_t)(fp & low_mask64(r));
  int homeBlockNum = (int)(fq / 64);
  int blockNum = homeBlockNum;
  uint32_t homeSlotNum = fq % 64;
  uint64_t occs = f.occ(blockNum);
  uint8_t blockOffset = f.offset(blockNum);
  uint32_t rank = orc_rank(occs, homeSlotNum);
  uint64_t runEnds = f.ends(blockNum);
  uint32_t endOfRun = orc_select(runEnds & ~low_mask64(blockOffset), rank);
  if (rank == 0) endOfRun = blockOffset == 0 ? 0 : (uint32_t)blockOffset - 1;
  while (endOfRun == ORC_NOT_FOUND) {
    rank -= (uint32_t)__builtin_popcountll(f.ends(blockNum) & ~low_mask64(blockOffset));
    if ((int)blockOffset - 64 > 0) blockOffset = (uint8_t)(blockOffset - 64);
    else blockOffset = 0;
    blockNum++;
    if ((uint32_t)blockNum >= f.NB) return ORC_NOT_FOUND;
    endOfRun = orc_select(f.ends(blockNum) & ~low_mask64(blockOffset), rank);
  }
  if (endOfRun >= 64) return ORC_NOT_FOUND;  // offset > 64 with rank 0: reference UB
  if (((uint64_t)blockNum * 64 + endOfRun < (uint64_t)homeBlockNum * 64 + homeSlotNum) |
      (blockOffset + rank == 0)) {
    f.set_ends(homeBlockNum, runEnds | (1ull << homeSlotNum));
    f.set_rem(homeBlockNum, homeSlotNum, fr);
    f.set_occ(homeBlockNum, occs | (1ull << homeSlotNum));
    return (uint32_t)homeBlockNum * 64 + homeSlotNum;
  }
  endOfRun++;
  if (endOfRun == 64) {
    endOfRun = 0;
    blockNum++;
    if ((uint32_t)blockNum >= f.NB) return ORC_NOT_FOUND;
  }
  int runEndBlock = blockNum;
  uint32_t unused = rsqf_first_unused(f, &blockNum, endOfRun);
  if (unused == ORC_NOT_FOUND || unused >= 64) return ORC_NOT_FOUND;
  if (blockNum > homeBlockNum)
    for (int i = 0; i < blockNum - homeBlockNum; ++i)
      f.set_offset(blockNum - i, (uint8_t)(f.offset(blockNum - i) + 1));
  while ((uint64_t)blockNum * 64 + unused > (uint64_t)runEndBlock * 64 + endOfRun) {
    if (unused == 0) {
      int nb = blockNum - 1;
      uint32_t ns = 63;
      f.set_rem(blockNum, unused, f.rem(nb, ns));
      if (bit_of(f.ends(nb), ns)) f.set_ends(blockNum, f.ends(blockNum) | 1ull);
      else f.set_ends(blockNum, f.ends(blockNum) & ~1ull);
      unused = 63;
      blockNum--;
    } else {
      f.set_rem(blockNum, unused, f.rem(blockNum, unused - 1));
      if (bit_of(f.ends(blockNum), unused - 1)) f.set_ends(blockNum, f.ends(blockNum) | (1ull << unused));
      else f.set_ends(blockNum, f.ends(blockNum) & ~(1ull << unused));
      unused--;
    }
  }
  if (!bit_of(f.occ(homeBlockNum), homeSlotNum)) {
    f.set_rem(blockNum, unused, fr);
    f.set_ends(blockNum, f.ends(blockNum) | (1ull << unused));
    f.set_occ(homeBlockNum, f.occ(homeBlockNum) | (1ull << homeSlotNum));
    return (uint32_t)blockNum * 64 + unused;
  }
  uint32_t nextSlot = unused - 1;
  int nextBlock = blockNum;
  if (unused == 0) {
    nextSlot = 63;
    nextBlock = blockNum - 1;
    if (nextBlock < 0) return ORC_NOT_FOUND;
  }
  f.set_ends(blockNum, f.ends(blockNum) | (1ull << unused));
  f.set_ends(nextBlock, f.ends(nextBlock) & ~(1ull << nextSlot));
  uint32_t nextRem = f.rem(nextBlock, nextSlot);
  do {
    if (nextRem <= fr) {
      f.set_rem(blockNum, unused, fr);
      return (uint32_t)blockNum * 64 + unused;
    }
    f.set_rem(blockNum, unused, nextRem);
    if (unused > 0) {
      unused--;
      if (unused == 0) {
        if (blockNum == 0) {
          f.set_rem(blockNum, unused, fr);
          return (uint32_t)blockNum * 64 + unused;
        }
        nextSlot = 63;
        nextBlock--;
      } else {
        nextSlot = unused - 1;
      }
    } else {
      unused = 63;
      blockNum--;
      if (blockNum < 0) return ORC_NOT_FOUND;
      nextSlot = unused - 1;
    }
    nextRem = f.rem(nextBlock, nextSlot);
  } while (!bit_of(f.ends(nextBlock), nextSlot) && ((uint64_t)nextBlock * 64 + nextSlot) >= fq);
  f.set_rem(blockNum, unused, fr);
  return (uint32_t)blockNum * 64 + unused;
}

void orc_rsqf_lookup_keys(const uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                          size_t n, uint32_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = orc_rsqf_lookup_fp(table, q, r, orc_fp_ap(keys[i], q, r));
}

size_t orc_rsqf_insert_keys(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                            size_t n, uint32_t* out) {
  size_t fails = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t s = orc_rsqf_insert_fp(table, q, r, orc_fp_ap(keys[i], q, r));
    if (out) out[i] = s;
    fails += (s == ORC_NOT_FOUND);
  }
  return fails;
}

// Decode of a canonical RSQF table: the k-th set occupieds bit (global order)
// owns the k-th run end; its run spans [max(fq_k, end_{k-1}+1), end_k].
size_t orc_rsqf_extract(const uint8_t* table, unsigned q, unsigned r, uint64_t* out) {
  Rsqf f = mk_rsqf(const_cast<uint8_t*>(table), q, r);
  uint64_t nslots = (uint64_t)f.NB * 64;
  uint64_t qo = 0, re = 0;  // scan positions for occupieds / runEnds
  int64_t prevEnd = -1;
  size_t k = 0;
  for (;;) {
    while (qo < nslots && !bit_of(f.occ((uint32_t)(qo / 64)), (uint32_t)(qo % 64))) qo++;
    while (re < nslots && !bit_of(f.ends((uint32_t)(re / 64)), (uint32_t)(re % 64))) re++;
    if (qo >= nslots || re >= nslots) break;
    int64_t start = std::max<int64_t>((int64_t)qo, prevEnd + 1);
    for (int64_t s = start; s <= (int64_t)re; ++s)
      out[k++] = (qo << r) | f.rem((uint32_t)(s / 64), (uint32_t)(s % 64));
    prevEnd = (int64_t)re;
    qo++;
    re++;
  }
  return k;
}

// ================================================================== CPU baseline loops
// Sequential loops timed like CPUInsertPerf.cpp:41-48 / CPULookupPerf.cpp:51-58
// (one call per key, monotonic clock around the whole loop).
static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double orc_time_sqf_insert(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys, size_t n) {
  double t0 = now_s();
  for (size_t i = 0; i < n; ++i) {
    int st;
    orc_sqf_insert_fp(table, q, r, orc_fp_ap(keys[i], q, r), &st);
  }
  return now_s() - t0;
}

double orc_time_sqf_lookup(const uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                           size_t n, uint32_t* out, int threads) {
  if (threads < 1) threads = 1;
  double t0 = now_s();
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    ts.emplace_back([=]() {
      size_t lo = n * t / threads, hi = n * (t + 1) / threads;
      for (size_t i = lo; i < hi; ++i) out[i] = orc_sqf_lookup_fp(table, q, r, orc_fp_ap(keys[i], q, r));
    });
  }
  for (auto& th : ts) th.join();
  return now_s() - t0;
}

double orc_time_sqf_delete(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys, size_t n) {
  double t0 = now_s();
  for (size_t i = 0; i < n; ++i) orc_sqf_delete_fp(table, q, r, orc_fp_ap(keys[i], q, r));
  return now_s() - t0;
}

double orc_time_rsqf_insert(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys, size_t n) {
  double t0 = now_s();
  for (size_t i = 0; i < n; ++i) orc_rsqf_insert_fp(table, q, r, orc_fp_ap(keys[i], q, r));
  return now_s() - t0;
}

double orc_time_rsqf_lookup(const uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                            size_t n, uint32_t* out, int threads) {
  if (threads < 1) threads = 1;
  double t0 = now_s();
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    ts.emplace_back([=]() {
      size_t lo = n * t / threads, hi = n * (t + 1) / threads;
      for (size_t i = lo; i < hi; ++i) out[i] = orc_rsqf_lookup_fp(table, q, r, orc_fp_ap(keys[i], q, r));
    });
  }
  for (auto& th : ts) th.join();
  return now_s() - t0;
}

}  // extern "C"
