// gqf_sqf_dev.hpp -- device code of the standard quotient filter (slot
// access, lookups, the fused rebuild, decode), shared by the per-filter
// entry points (gqf_sqf.hip) and the batched ones (gqf_batch.hip).
#pragma once

#include <type_traits>

#include "gqf_rebuild.hpp"

namespace gqf {

constexpr uint32_t SQF_NOT_FOUND = 0xFFFFFFFFu;

// numElements (quotientFilter.cu:174): (1 << q) * 1.1 in double, truncated.
static inline uint32_t sqf_slots(unsigned q) { return (uint32_t)((double)(1u << q) * 1.1); }

// Slot widths with (r+3) % 8 != 0 (the reference GPU code requires
// (r+3) % 8 == 0, README.md:32-33): the same S slots and the same elements,
// (r+3) bits each, packed MSB-first into 32-bit words in the CPU_version's
// order (getElement, quotientFilter.cpp:114-130: slot i at bits
// [i(r+3), (i+1)(r+3)) counted from the most significant bit of word 0),
// plus one trailing word so a 64-bit read at the last slot stays inside.
// Such a filter has bytesPerElement == 0.  Supported up to 16-bit elements.
__host__ __device__ __forceinline__ bool sqf_is_packed(unsigned r) { return (r + 3) % 8 != 0; }
static inline size_t sqf_packed_words(unsigned q, unsigned r) {
  return ((size_t)sqf_slots(q) * (r + 3) + 31) / 32 + 1;
}
// calcNumSlotsGPU (quotientFilter.cu:37-42): bytes; packed widths: the words.
static inline size_t sqf_bytes(unsigned q, unsigned r) {
  if (sqf_is_packed(r)) return 4 * sqf_packed_words(q, r);
  size_t tableBits = (size_t)(1u << q) * (r + 3);
  return (size_t)((double)(tableBits / 8) * 1.1);
}
static inline bool sqf_geometry_ok(unsigned q, unsigned r, unsigned bpe) {
  if (q < 1 || r < 1 || q + r > 32 || q > 28) return false;
  return sqf_is_packed(r) ? (r + 3 <= 16 && bpe == 0) : bpe == (r + 3) / 8;
}

[[maybe_unused]] static gqf_status sqf_check(const gqf_sqf& qf) {
  if (!qf.table) {
    set_error("SQF table is NULL (call gqf_sqf_init)");
    return GQF_ERR_INVALID;
  }
  if (!sqf_geometry_ok(qf.qbits, qf.rbits, qf.bytesPerElement)) {
    set_error("unsupported SQF geometry q=%u r=%u bytesPerElement=%u (need q+r<=32, q<=28, r+3<=16 for packed "
              "widths; bytesPerElement (r+3)/8, or 0 when (r+3)%%8 != 0)",
              qf.qbits, qf.rbits, qf.bytesPerElement);
    return GQF_ERR_INVALID;
  }
  return GQF_OK;
}

// Element of slot i of a packed table of `width`-bit elements (0 past S).
__host__ __device__ __forceinline__ uint32_t sqf_get_packed(const uint32_t* __restrict__ w, uint32_t S,
                                                           uint32_t width, uint32_t i) {
  if (i >= S) return 0u;
  const uint64_t bit = (uint64_t)i * width;
  const uint64_t j = bit >> 5;
  const uint32_t sh = (uint32_t)(bit & 31u);
  const uint64_t win = ((uint64_t)w[j] << 32) | w[j + 1];
  return (uint32_t)(win >> (64u - sh - width)) & ((1u << width) - 1u);
}

// Packed table -> wide table (16-bit big-endian elements, the layout of a
// byte-aligned 2-byte table), one slot per thread.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_sqf_unpack(const uint32_t* __restrict__ packed,
                                                                          uint32_t S, uint32_t width,
                                                                          uint16_t* __restrict__ wide) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= S) return;
  const uint32_t e = sqf_get_packed(packed, S, width, i);
  wide[i] = (uint16_t)((e >> 8) | ((e & 0xFFu) << 8));
}

// Wide table -> packed words, one word per thread: the slots whose bits
// overlap word j, shifted into place.  With ctr, nothing is written when the
// call's capacity flag is set (the wide table then holds no result).
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_sqf_pack(const uint16_t* __restrict__ wide,
                                                                        uint32_t S, uint32_t width,
                                                                        uint32_t* __restrict__ packed,
                                                                        uint32_t nwords,
                                                                        const uint32_t* __restrict__ ctr) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= nwords) return;
  if (ctr && (ctr[CTR_ERR] & DEV_ERR_CAPACITY)) return;
  const uint64_t b0 = (uint64_t)j * 32;
  uint32_t acc = 0;
  for (uint64_t s = b0 / width; s < S && s * width < b0 + 32; ++s) {
    const uint32_t x = wide[s];
    const uint32_t e = ((x & 0xFFu) << 8) | (x >> 8);
    // the element's most significant bit sits at bit offset s*width - b0 of the word
    const int sh = 32 - (int)width - (int)((int64_t)(s * width) - (int64_t)b0);
    acc |= sh >= 0 ? (e << sh) : (e >> -sh);
  }
  packed[j] = acc;
}

// ---------------------------------------------------------------- slot access
template <int BPE>
__host__ __device__ __forceinline__ uint32_t sqf_get(const uint8_t* __restrict__ t, uint32_t S, uint32_t i) {
  if (i >= S) return 0u;  // past the last slot reads as empty
  if (BPE == 1) return t[i];
  const uint8_t* p = t + (size_t)i * BPE;
  uint32_t e = p[0];
#pragma unroll
  for (int k = 1; k < BPE; ++k) e = (e << 8) | p[k];
  return e;
}

// Returned by a walk that ran out of its budget (never a slot index: slots
// are below 2^29).
constexpr uint32_t SQF_GIVE_UP = 0xFFFFFFFEu;

// findRunStartGPU (quotientFilter.cu:172-196) + run scan of hashAndLookUp
// (quotientFilter.cu:333-356): slot of the first matching remainder, over any
// element accessor get(i) (0 past the last slot).
// budget: at most 64 * budget slot steps, then SQF_GIVE_UP.
template <class Get>
__host__ __device__ __forceinline__ uint32_t sqf_walk(const Get& get, uint32_t S, uint32_t fq, uint32_t fr,
                                                     uint32_t budget) {
  if (!(get(fq) & 4u)) return SQF_NOT_FOUND;
  const uint64_t lim = (uint64_t)budget * 64;
  uint64_t steps = 0;
  uint32_t b = fq;
  while (b > 0 && (get(b) & 1u)) {
    b--;
    if (++steps > lim) return SQF_GIVE_UP;
  }
  uint32_t s = b;
  while (b != fq) {
    do {
      s++;
      if (++steps > lim) return SQF_GIVE_UP;
    } while ((get(s) & 2u) && s < S);
    do {
      b++;
      if (++steps > lim) return SQF_GIVE_UP;
    } while (!(get(b) & 4u) && b < S);
  }
  uint32_t e = get(s);
  do {
    const uint32_t rem = e >> 3;
    if (rem == fr) return s;
    if (rem > fr) return SQF_NOT_FOUND;
    s++;
    e = get(s);
    if (++steps > lim) return SQF_GIVE_UP;
  } while (e & 2u);
  return SQF_NOT_FOUND;
}

template <int BPE>
struct SqfBytes {
  const uint8_t* t;
  uint32_t S;
  __host__ __device__ __forceinline__ uint32_t operator()(uint32_t i) const { return sqf_get<BPE>(t, S, i); }
};
struct SqfPacked {
  const uint32_t* w;
  uint32_t S, width;
  __host__ __device__ __forceinline__ uint32_t operator()(uint32_t i) const { return sqf_get_packed(w, S, width, i); }
};

template <int BPE>
__host__ __device__ __forceinline__ uint32_t sqf_lookup_one(const uint8_t* __restrict__ t, uint32_t S, uint32_t fq,
                                                           uint32_t fr, uint32_t budget = 0xFFFFFFFFu) {
  return sqf_walk(SqfBytes<BPE>{t, S}, S, fq, fr, budget);
}

// The same walk over 64-slot windows (BPE 1 and 2): a window is loaded with
// 16-byte loads (64 or 128 table bytes, 64-slot aligned) and turned into
// occupied / continuation / shifted bit masks, so that the cluster start, the
// k-th run start and the first remainder >= fr are found with popcount and
// count-trailing-zeros; the walk moves one window back or forward only when
// it crosses a window edge.  Slots at or past S read as empty, as above.
template <int BPE>
struct SqfWin {
  uint32_t raw[16 * BPE];
  uint64_t occ, cont, shf;
};

// Bit 0 of each byte of a, then of b (other bits zero) -> bits 0..7, one
// multiply: with b's bits moved to bit 4 of each byte, the product's partial
// terms (byte j, bit t, multiplier bit s in {3, 10, 17, 24}) sit at distinct
// positions 8j + t + s, so no carries occur, and bits 24..31 hold exactly
// a's bits (t = 0) then b's (t = 4).
__host__ __device__ __forceinline__ uint32_t byte_bits8(uint32_t a, uint32_t b) {
  return ((a | (b << 4)) * 0x01020408u) >> 24;
}

template <int BPE>
__host__ __device__ __forceinline__ void sqf_win_load(const uint8_t* __restrict__ t, uint32_t S, uint32_t w, SqfWin<BPE>& W) {
  const uint64_t s0 = (uint64_t)w * 64;
  if (s0 + 64 <= S) {
    const uint4* p = reinterpret_cast<const uint4*>(t + s0 * BPE);
#pragma unroll
    for (int k = 0; k < 4 * BPE; ++k) {
      const uint4 x = p[k];
      W.raw[4 * k] = x.x;
      W.raw[4 * k + 1] = x.y;
      W.raw[4 * k + 2] = x.z;
      W.raw[4 * k + 3] = x.w;
    }
  } else {
    // the table's last window (slots past S read as empty): 4-byte loads of
    // the words inside the table, byte loads for the one word that crosses
    // its end
    const uint64_t end = (uint64_t)S * BPE;
#pragma unroll
    for (int k = 0; k < 16 * BPE; ++k) {
      const uint64_t at = s0 * BPE + 4 * (uint64_t)k;
      uint32_t v = 0;
      if (at + 4 <= end) v = *reinterpret_cast<const uint32_t*>(t + at);
      else
        for (uint64_t b = at; b < end; ++b) v |= (uint32_t)t[b] << (8 * (uint32_t)(b - at));
      W.raw[k] = v;
    }
  }
  uint64_t occ = 0, cont = 0, shf = 0;
  if (BPE == 1) {
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
      const uint32_t x = W.raw[k], y = W.raw[k + 1];
      occ |= (uint64_t)byte_bits8((x >> 2) & 0x01010101u, (y >> 2) & 0x01010101u) << (4 * k);
      cont |= (uint64_t)byte_bits8((x >> 1) & 0x01010101u, (y >> 1) & 0x01010101u) << (4 * k);
      shf |= (uint64_t)byte_bits8(x & 0x01010101u, y & 0x01010101u) << (4 * k);
    }
  } else {
    // slot j = bytes 2j (high) and 2j+1 (low, holding the flags)
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const uint32_t x = W.raw[k];
      const uint32_t f = ((x >> 8) & 7u) | ((x >> 21) & 0x38u);  // flags of slot 2k in bits 0..2, 2k+1 in 3..5
      occ |= (uint64_t)(((f >> 2) & 1u) | ((f >> 4) & 2u)) << (2 * k);
      cont |= (uint64_t)(((f >> 1) & 1u) | ((f >> 3) & 2u)) << (2 * k);
      shf |= (uint64_t)((f & 1u) | ((f >> 2) & 2u)) << (2 * k);
    }
  }
  W.occ = occ;
  W.cont = cont;
  W.shf = shf;
}

// Bit j set iff remainder(slot j of the window) >= x  (x <= 2^r).
template <int BPE>
__host__ __device__ __forceinline__ uint64_t sqf_win_ge(const SqfWin<BPE>& W, uint32_t x) {
  uint64_t m = 0;
  if (BPE == 1) {
    // per byte: (rem | 0x80) - x keeps bit 7 iff rem >= x (rem, x < 128: no borrow between bytes)
    const uint32_t xr = x * 0x01010101u;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
      const uint32_t a = (W.raw[k] >> 3) & 0x1F1F1F1Fu, b = (W.raw[k + 1] >> 3) & 0x1F1F1F1Fu;
      const uint32_t ga = ((a | 0x80808080u) - xr) & 0x80808080u;
      const uint32_t gb = ((b | 0x80808080u) - xr) & 0x80808080u;
      m |= (uint64_t)byte_bits8(ga >> 7, gb >> 7) << (4 * k);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const uint32_t x0 = W.raw[k];
      const uint32_t r0 = (((x0 & 0xFFu) << 8) | ((x0 >> 8) & 0xFFu)) >> 3;
      const uint32_t r1 = ((((x0 >> 16) & 0xFFu) << 8) | (x0 >> 24)) >> 3;
      m |= (uint64_t)((r0 >= x ? 1u : 0u) | (r1 >= x ? 2u : 0u)) << (2 * k);
    }
  }
  return m;
}

// budget: at most that many window loads after the first, then SQF_GIVE_UP.
template <int BPE>
__host__ __device__ __forceinline__ uint32_t sqf_lookup_win(const uint8_t* __restrict__ t, uint32_t S, uint32_t fq,
                                                   uint32_t fr, uint32_t budget = 0xFFFFFFFFu) {
  uint32_t loads = 0;
  SqfWin<BPE> W;
  uint32_t cw = fq >> 6;
  const uint32_t rel = fq & 63u;
  sqf_win_load<BPE>(t, S, cw, W);
  if (!((W.occ >> rel) & 1ull)) return SQF_NOT_FOUND;
  // cluster start b: the last non-shifted slot at or before fq (or slot 0);
  // k = occupied slots in (b, fq] = run starts to skip after b
  uint64_t lim = rel == 63 ? ~0ull : ((2ull << rel) - 1ull);  // bits [0, rel]
  uint32_t k = 0, brel = 0;
  for (;;) {
    const uint64_t ns = ~W.shf & lim;
    if (ns) {
      brel = 63u - (uint32_t)__builtin_clzll(ns);
      k += (uint32_t)__builtin_popcountll(W.occ & lim & ~((2ull << brel) - 1ull));
      break;
    }
    k += (uint32_t)__builtin_popcountll(W.occ & lim);
    if (cw == 0) {
      brel = 0;
      k -= (uint32_t)(W.occ & 1ull);
      break;
    }
    if (++loads > budget) return SQF_GIVE_UP;
    --cw;
    sqf_win_load<BPE>(t, S, cw, W);
    lim = ~0ull;
  }
  // s: the k-th slot after b without a continuation bit
  uint32_t srel = brel;
  uint64_t from = brel == 63 ? 0ull : (~0ull << (brel + 1));
  while (k > 0) {
    uint64_t rs = ~W.cont & from;
    const uint32_t c = (uint32_t)__builtin_popcountll(rs);
    if (c >= k) {
      for (uint32_t i = 1; i < k; ++i) rs &= rs - 1ull;
      srel = (uint32_t)__builtin_ctzll(rs);
      break;
    }
    k -= c;
    if (++loads > budget) return SQF_GIVE_UP;
    ++cw;
    sqf_win_load<BPE>(t, S, cw, W);
    from = ~0ull;
  }
  // the run: s and the continuation slots after it; first remainder >= fr
  uint64_t run_from = 1ull << srel;
  uint64_t after = srel == 63 ? 0ull : (~0ull << (srel + 1));
  for (;;) {
    const uint64_t stop = ~W.cont & after;
    const uint64_t run = (run_from | after) & (stop ? ((1ull << __builtin_ctzll(stop)) - 1ull) : ~0ull);
    const uint64_t ge = sqf_win_ge<BPE>(W, fr) & run;
    if (ge) {
      const uint32_t p = (uint32_t)__builtin_ctzll(ge);
      return ((sqf_win_ge<BPE>(W, fr + 1) >> p) & 1ull) ? SQF_NOT_FOUND : cw * 64u + p;
    }
    if (stop) return SQF_NOT_FOUND;
    if (++loads > budget) return SQF_GIVE_UP;
    ++cw;
    sqf_win_load<BPE>(t, S, cw, W);
    run_from = 0;
    after = ~0ull;
  }
}

template <int BPE, bool WIN>
__device__ __forceinline__ uint32_t sqf_lookup_fp(const uint8_t* __restrict__ t, uint32_t S, uint32_t fq,
                                                  uint32_t fr, uint32_t budget = 0xFFFFFFFFu) {
  if (WIN) return sqf_lookup_win<BPE>(t, S, fq, fr, budget);
  return sqf_lookup_one<BPE>(t, S, fq, fr, budget);
}

// Slot of the k-th run start (1-based, counting from slot 0) through the
// window index, searching windows from vmin on; S when the table holds fewer
// than k runs (or the index disagrees with the table).
template <int BPE>
__device__ uint32_t sqf_select_run(const uint8_t* __restrict__ t, uint32_t S, const uint2* __restrict__ idx,
                                   uint32_t k, uint32_t vmin) {
  const uint32_t nwin = (S + 63) / 64;
  if (k == 0 || idx[nwin].y < k || idx[vmin].y >= k) return S;
  uint32_t lo = vmin, hi = nwin;  // idx[lo].y < k <= idx[hi].y
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (idx[mid].y < k) lo = mid;
    else hi = mid;
  }
  uint32_t need = k - idx[lo].y;
  const uint32_t end = lo * 64 + 64 < S ? lo * 64 + 64 : S;
  for (uint32_t s = lo * 64; s < end; ++s) {
    const uint32_t e = sqf_get<BPE>(t, S, s);
    if ((e & 7u) != 0 && !(e & 2u) && --need == 0) return s;  // occupied and not a continuation
  }
  return S;
}

// The same result as the walk, through a window index, for walks through
// long clusters.  idx[w] = {occupied quotients, run starts} in the slots
// before 64 w, for w = 0 .. ceil(S / 64) (the last entry: the totals).  In a
// canonical table the k-th run start (from slot 0) begins the run of the
// k-th occupied quotient, so fq's run [s, N) is bracketed by two index
// searches (N: the next run start).  Over (s, N] the predicate "no
// continuation bit, or remainder >= fr" is monotone (the run's remainders
// ascend, and only empty slots lie between a run's end and the next run), so
// the first remainder >= fr is found by a binary search.  fq must be
// occupied.
template <int BPE>
__device__ uint32_t sqf_lookup_indexed(const uint8_t* __restrict__ t, uint32_t S, const uint2* __restrict__ idx,
                                       uint32_t fq, uint32_t fr) {
  const uint32_t w = fq >> 6;
  uint32_t k = idx[w].x;  // 1-based rank of fq among the occupied quotients
  for (uint32_t s = w * 64; s <= fq; ++s) k += (sqf_get<BPE>(t, S, s) >> 2) & 1u;
  // a run of a quotient >= 64 w starts at or after slot 64 w
  const uint32_t s = sqf_select_run<BPE>(t, S, idx, k, w);
  if (s >= S) return SQF_NOT_FOUND;
  const uint32_t e0 = sqf_get<BPE>(t, S, s);
  if ((e0 >> 3) >= fr) return (e0 >> 3) == fr ? s : SQF_NOT_FOUND;
  uint32_t a = s, b = sqf_select_run<BPE>(t, S, idx, k + 1, s >> 6);  // P(a) false, P(b) true
  while (b - a > 1) {
    const uint32_t m = a + (b - a) / 2;
    const uint32_t e = sqf_get<BPE>(t, S, m);
    if (!(e & 2u) || (e >> 3) >= fr) b = m;
    else a = m;
  }
  const uint32_t e = sqf_get<BPE>(t, S, b);  // slot S reads as empty
  return ((e & 2u) && (e >> 3) == fr) ? b : SQF_NOT_FOUND;
}

template <int BPE, bool WIN>
__global__ __launch_bounds__(256) void k_sqf_lookup(const uint8_t* __restrict__ t, uint32_t S, uint32_t q, uint32_t r,
                                                    const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t* __restrict__ out, uint32_t hflag) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint32_t fp = fingerprint32(keys[i], q + r + hflag);
  out[i] = sqf_lookup_fp<BPE, WIN>(t, S, (fp >> r) & (uint32_t)mask64(q), fp & ((1u << r) - 1u));
}

// Lookups on a packed table (slot widths with (r+3) % 8 != 0): the same walk.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_sqf_lookup_packed(
    const uint32_t* __restrict__ t, uint32_t S, uint32_t q, uint32_t r, const uint32_t* __restrict__ keys, uint32_t n,
    uint32_t* __restrict__ out, uint32_t hflag) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint32_t fp = fingerprint32(keys[i], q + r + hflag);
  out[i] = sqf_walk(SqfPacked{t, S, r + 3}, S, (fp >> r) & (uint32_t)mask64(q), fp & ((1u << r) - 1u), 0xFFFFFFFFu);
}

// ---------------------------------------------------------------- fused rebuild
// From sorted fingerprints to the finished table: optional merge of two sorted
// inputs (merge path, A first on ties), the keep rule (all / unique /
// dedup-insert / delete flags), the canonical locations
// loc_i = max(fq_i, loc_{i-1}+1), and the table bytes.  A block of entries
// acts on the incoming location as x -> max(x + c, l); the states (c, l, f)
// (f: quotient of the last kept entry, which decides run heads) are combined
// in chunk order:
//   k_rb_agg    each chunk's state (one read of the inputs),
//   rb_scan     exclusive scan over the chunks (two small launches); the total
//               decides capacity before anything is written,
//   k_rb_write  each chunk walks its 4096-entry sub-tiles with a running
//               prefix; sub-tile u owns the slot range (L_{u-1}, L_u] ending at
//               its last entry and writes every byte of it through an LDS
//               window, each thread the slots up to its own last entry from
//               registers; slots past the last entry are cleared by
//               k_sqf_clear_tail.  An occupied bit that falls in an earlier
//               sub-tile's range goes to a fix-up list (k_sqf_fixup).
// Replaces the location and write stages of the three reference builds
// (quotientFilter.cu:788-857, 985-1038, 1492-1539), the merge rebuild
// (quotientFilter.cu:1651-1719) and the end states of insert /
// superclusterDeletes.
constexpr uint32_t RB_MERGE_PER = 2;  // sub-tiles per workgroup in merge rebuilds (staged ahead)
constexpr uint32_t RB_FIX = 64;       // per-sub-tile occupied-bit fix-up slots
// Slots per write window: 8192, 4096 for 3- and 4-byte elements (the
// window and the two merge stages then fit three workgroups per CU).
template <int BPE>
constexpr uint32_t rb_win_slots() {
  return BPE <= 2 ? 8192u : 4096u;
}
// LDS words of a write window: its table bytes plus 16, so that window word 0
// sits on a 16-byte boundary of the table.
template <int BPE>
constexpr uint32_t rb_win_words() {
  return (rb_win_slots<BPE>() * BPE + 16) / 4;
}

// This thread's RB_ITEMS consecutive entries of sub-tile [d0, d0+tn) of A
// (no merge): v[1+j] = entry j and v[0] = the entry before them; fl =
// delete flags of the entries (KEEP_FLAGS); 16-byte loads.  Merge rebuilds
// stage their inputs instead (rb_stage_issue / rb_merge_staged).
template <int MODE>
__device__ __forceinline__ void rb_load(const uint32_t* __restrict__ A, const uint8_t* __restrict__ flags,
                                        uint32_t d0, uint32_t tn, uint32_t* v, uint32_t& tags, uint32_t& fl) {
  const int tid = threadIdx.x;
  const uint32_t i0 = (uint32_t)tid * RB_ITEMS;
  fl = 0;
  tags = 0;
  {
    const uint64_t g0 = (uint64_t)d0 + i0;
    uint32_t w[RB_ITEMS];
    // 16-byte loads need A + g0 16-byte aligned: d0 and i0 are multiples of
    // 8 entries, so that holds for every thread when it holds for A + d0
    // (plain tables: A is 256-byte aligned; batches: a filter's segment of
    // the batch key array may start at any entry)
    if (i0 + RB_ITEMS <= tn && (reinterpret_cast<uintptr_t>(A + d0) & 15u) == 0) {
      const uint4* p = reinterpret_cast<const uint4*>(A + g0);
#pragma unroll
      for (int k = 0; k < RB_ITEMS / 4; ++k) {
        const uint4 x = p[k];
        w[4 * k] = x.x;
        w[4 * k + 1] = x.y;
        w[4 * k + 2] = x.z;
        w[4 * k + 3] = x.w;
      }
      if (MODE == KEEP_FLAGS) {
        const uint2 f = *reinterpret_cast<const uint2*>(flags + g0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          fl |= ((f.x >> (8 * k)) & 1u) << k;
          fl |= ((f.y >> (8 * k)) & 1u) << (4 + k);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < RB_ITEMS; ++k) {
        w[k] = (i0 + k < tn) ? A[g0 + k] : 0u;
        if (MODE == KEEP_FLAGS && i0 + k < tn && flags[g0 + k]) fl |= 1u << k;
      }
    }
    v[0] = g0 > 0 && i0 <= tn ? A[g0 - 1] : 0u;
#pragma unroll
    for (int k = 0; k < RB_ITEMS; ++k) v[1 + k] = w[k];
  }
}

// Double-buffered staging of the merge inputs (insert / merge rebuilds):
// sub-tile u's A range [a0, a1) and B range are copied into an LDS buffer
// with global->LDS loads, which take no VGPRs and run on while the
// workgroup works on the previous sub-tile; the entries just before the
// ranges (the first thread's previous entry) go to prev2[0] (A) and
// prev2[1] (B).  rb_stage_wait() and a barrier make a buffer readable.
// FULL: tn == RB_TILE (every position is in range, no per-entry tests).
template <bool FULL = false>
__device__ __forceinline__ void rb_stage_issue(const uint32_t* __restrict__ A, const uint32_t* __restrict__ B,
                                               uint32_t a0, uint32_t a1, uint32_t d0, uint32_t tn, uint32_t* buf,
                                               uint32_t* prev2) {
  const uint32_t la = a1 - a0, b0 = d0 - a0;
  // wid is uniform in the wave: scalar, so the LDS destinations are too
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // this lane's sources for step 0, as addresses (the B one is used only
  // where pos >= la); each step moves RB_WAVES * 64 positions on, so the
  // steps need no table bases
  const uint32_t pos0 = wid * 64 + lane;
  const uintptr_t pa = reinterpret_cast<uintptr_t>(A) + 4 * ((uintptr_t)a0 + pos0);
  const uintptr_t pb = reinterpret_cast<uintptr_t>(B) + 4 * ((uintptr_t)b0 + pos0) - 4 * (uintptr_t)la;
#pragma unroll
  for (int k = 0; k < RB_ITEMS; ++k) {
    // wave wid, step k: the 64 consecutive positions from base (lane l -> buf[base + l])
    const uint32_t base = ((uint32_t)k * RB_WAVES + wid) * 64;
    const uint32_t pos = base + lane;
    if (FULL || pos < tn) {
      const uintptr_t step = (uintptr_t)k * RB_WAVES * 64 * 4;
      const uint32_t* src = reinterpret_cast<const uint32_t*>((pos < la ? pa : pb) + step);
      __builtin_amdgcn_global_load_lds(src, buf + base, 4, 0, 0);
    }
  }
  if (wid == 0 && lane < 2 && (lane == 0 ? a0 > 0 : b0 > 0))
    __builtin_amdgcn_global_load_lds(lane == 0 ? A + (a0 - 1) : B + (b0 - 1), prev2, 4, 0, 0);
}

// rb_stage_issue, with the whole-tile form for whole tiles.
__device__ __forceinline__ void rb_stage(const uint32_t* __restrict__ A, const uint32_t* __restrict__ B, uint32_t a0,
                                         uint32_t a1, uint32_t d0, uint32_t tn, uint32_t* buf, uint32_t* prev2) {
  if (tn == RB_TILE) rb_stage_issue<true>(A, B, a0, a1, d0, tn, buf, prev2);
  else rb_stage_issue<false>(A, B, a0, a1, d0, tn, buf, prev2);
}

// This thread's RB_ITEMS merged entries of a staged sub-tile (as rb_load;
// tags bit j: entry j came from B).  A batch adds few entries to a sub-tile,
// so instead of a merge-path search per thread every B entry finds its
// output position (its index + the A entries <= it: A first on ties) and
// marks it in the bitmap s_bmp (RB_TILE bits, zero on entry); a prefix count
// of the bitmap words (s_wpre) gives each thread its A and B offsets.  Every
// thread of the block calls it (two barriers inside).  FULL: tn == RB_TILE.
template <bool FULL = false>
__device__ __forceinline__ void rb_merge_staged(const uint32_t* buf, const uint32_t* prev2, uint32_t a0, uint32_t a1,
                                                uint32_t d0, uint32_t tn, uint32_t* v, uint32_t& tags,
                                                uint32_t* s_bmp, uint32_t* s_wpre) {
  const uint32_t b0 = d0 - a0, la = a1 - a0, lb = tn - la;
  const uint32_t* sA = buf;
  const uint32_t* sB = buf + la;
  for (uint32_t j = threadIdx.x; j < lb; j += RB_THREADS) {
    const uint32_t x = sB[j];
    uint32_t lo = 0, hi = la;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (sA[mid] <= x) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t p = j + lo;
    atomicOr(&s_bmp[p >> 5], 1u << (p & 31));
  }
  __syncthreads();
  static_assert(RB_TILE == 64 * 64, "one wave counts the bitmap, two words per lane");
  if (threadIdx.x < 64) {
    const uint32_t l = threadIdx.x;
    const uint32_t c0 = (uint32_t)__popc(s_bmp[2 * l]), c1 = (uint32_t)__popc(s_bmp[2 * l + 1]);
    const uint32_t inc = wave_incl_add(c0 + c1);
    s_wpre[2 * l] = inc - c0 - c1;
    s_wpre[2 * l + 1] = inc - c1;
  }
  __syncthreads();
  const uint32_t i0 = (uint32_t)threadIdx.x * RB_ITEMS;  // RB_ITEMS positions inside one bitmap word
  const uint32_t w = s_bmp[i0 >> 5];
  uint32_t bi = s_wpre[i0 >> 5] + (uint32_t)__popc(w & ((1u << (i0 & 31)) - 1u));
  uint32_t ai = i0 - bi;  // bi counts marked positions below i0
  // v[0]: the entry before position i0 (the sub-tile's predecessor for i0 = 0)
  if (i0 == 0) {
    const bool hasA = a0 > 0, hasB = b0 > 0;
    const uint32_t pA = hasA ? prev2[0] : 0u, pB = hasB ? prev2[1] : 0u;
    v[0] = (hasB && (!hasA || pA <= pB)) ? pB : pA;
  } else if (FULL || i0 <= tn) {
    const bool pb = (s_bmp[(i0 - 1) >> 5] >> ((i0 - 1) & 31)) & 1u;
    v[0] = pb ? sB[bi - 1] : sA[ai - 1];
  } else {
    v[0] = 0u;
  }
  const uint32_t bits = w >> (i0 & 31);
  // branch-free: one LDS read per entry, from B (buf + la) or A
  tags = FULL || i0 + RB_ITEMS <= tn ? bits & ((1u << RB_ITEMS) - 1u)
                             : (i0 < tn ? bits & ((1u << (tn - i0)) - 1u) : 0u);
#pragma unroll
  for (int j = 0; j < RB_ITEMS; ++j) {
    const uint32_t isb = (bits >> j) & 1u;
    const uint32_t at = isb ? la + bi : ai;
    v[1 + j] = FULL || i0 + j < tn ? buf[at] : 0u;
    bi += isb;
    ai += isb ^ 1u;
  }
}

// Keep rule of this thread's entries (bit j of the result) and their scan
// state.  FULL: tn == RB_TILE.
template <int MODE, bool FULL = false>
__device__ __forceinline__ uint32_t rb_classify(const uint32_t* v, uint32_t tags, uint32_t fl, uint32_t d0,
                                                uint32_t tn, uint32_t r, uint64_t mq, RS& agg) {
  uint32_t keepbits = 0;
  agg = rs_none();
  const uint32_t i0 = (uint32_t)threadIdx.x * RB_ITEMS;
#pragma unroll
  for (int j = 0; j < RB_ITEMS; ++j) {
    const uint32_t i = i0 + j;
    if (FULL || i < tn) {
      const uint32_t fp = v[1 + j], pfp = v[j];
      const bool first = (d0 + i) == 0;
      bool kp;
      if (MODE == KEEP_ALL) kp = true;
      else if (MODE == KEEP_UNIQUE) kp = first || fp != pfp;
      else if (MODE == KEEP_TAGGED_INSERT) kp = ((tags >> j) & 1u) == 0 || first || fp != pfp;
      else kp = (fl >> j) & 1u;
      if (kp) {
        keepbits |= 1u << j;
        rs_push(agg, (uint32_t)(((uint64_t)fp >> r) & mq));
      }
    }
  }
  return keepbits;
}

template <int MODE, bool MERGE>
__global__ __launch_bounds__(RB_THREADS, 8) void k_rb_agg(const uint32_t* __restrict__ A, uint32_t na,
                                                const uint32_t* __restrict__ na_dev,
                                                const uint32_t* __restrict__ B, uint32_t nb,
                                                const uint8_t* __restrict__ flags, const uint32_t* __restrict__ splits,
                                                uint32_t q, uint32_t r, uint32_t chunk,
                                                uint4* __restrict__ agg, uint32_t* __restrict__ ctr) {
  // the rebuild's counters start at zero (nothing reads them before the scan)
  if (blockIdx.x == 0 && threadIdx.x >= CTR_OP && threadIdx.x < CTR_TOTAL) ctr[threadIdx.x] = 0u;
  __shared__ uint32_t s_stage[MERGE ? 2 : 1][MERGE ? RB_TILE : 1];
  __shared__ uint32_t s_prev[2][2];
  __shared__ uint32_t s_bmp[MERGE ? RB_TILE / 32 : 1], s_wpre[MERGE ? RB_TILE / 32 : 1];
  __shared__ uint4 s_wave[RB_WAVES + 1];
  if (na_dev) na = *na_dev;
  const uint64_t N = (uint64_t)na + (MERGE ? nb : 0u);
  const uint64_t mq = mask64(q);
  const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t c1 = c0 + chunk < N ? c0 + chunk : N;
  RS X = rs_none();
  // MERGE: the next sub-tile's inputs are staged while this one is processed
  // splits: ca* of the current sub-tile, na* of the next one (loaded a
  // sub-tile ahead, so that issuing its stage does not wait on them)
  uint32_t ca0 = 0, ca1 = 0, na0 = 0, na1 = 0;
  if (MERGE && c0 < c1) {
    const uint32_t u0 = (uint32_t)(c0 / RB_TILE);
    ca0 = splits[u0];
    ca1 = splits[u0 + 1];
    if (c0 + RB_TILE < c1) na1 = splits[u0 + 2];
    na0 = ca1;
    rb_stage(A, B, ca0, ca1, (uint32_t)c0, (uint32_t)(c1 - c0 < RB_TILE ? c1 - c0 : RB_TILE), s_stage[0],
                   s_prev[0]);
  }
  int buf = 0;
  for (uint64_t d0 = c0; d0 < c1; d0 += RB_TILE) {
    const uint32_t tn = (uint32_t)(c1 - d0 < (uint64_t)RB_TILE ? c1 - d0 : (uint64_t)RB_TILE);
    uint32_t v[RB_ITEMS + 1];
    uint32_t tags, fl = 0;
    if (MERGE) {
      const uint32_t u = (uint32_t)(d0 / RB_TILE);
      const uint64_t dn = d0 + RB_TILE;
      rb_stage_wait();
      // the previous merge's bitmap reads ended before the barriers since
      if (threadIdx.x < RB_TILE / 32) s_bmp[threadIdx.x] = 0u;
      __syncthreads();
      if (dn < c1)
        rb_stage(A, B, na0, na1, (uint32_t)dn, (uint32_t)(c1 - dn < RB_TILE ? c1 - dn : RB_TILE),
                       s_stage[buf ^ 1], s_prev[buf ^ 1]);
      const uint32_t nn1 = dn + RB_TILE < c1 ? splits[u + 3] : 0u;
      if (tn == RB_TILE)
        rb_merge_staged<true>(s_stage[buf], s_prev[buf], ca0, ca1, (uint32_t)d0, tn, v, tags, s_bmp, s_wpre);
      else
        rb_merge_staged<false>(s_stage[buf], s_prev[buf], ca0, ca1, (uint32_t)d0, tn, v, tags, s_bmp, s_wpre);
      ca0 = na0;
      ca1 = na1;
      na0 = na1;
      na1 = nn1;
      buf ^= 1;
    } else {
      rb_load<MODE>(A, flags, (uint32_t)d0, tn, v, tags, fl);
    }
    RS a;
    if (tn == RB_TILE) (void)rb_classify<MODE, true>(v, tags, fl, (uint32_t)d0, tn, r, mq, a);
    else (void)rb_classify<MODE, false>(v, tags, fl, (uint32_t)d0, tn, r, mq, a);
    RS ex;
    const RS sub = rs_block(a, &ex, s_wave);
    X = rs_combine(X, sub);
    __syncthreads();
  }
  if (threadIdx.x == 0) agg[blockIdx.x] = rs_pack(X);
}

// ORs element e (BPE bytes, big-endian) into the window at byte offset off.
template <int BPE>
__device__ __forceinline__ void win_or(uint32_t* s_win, uint32_t off, uint32_t e) {
  uint32_t m;  // the element's bytes in memory order, as a little-endian word
  if (BPE == 1) m = e;
  else if (BPE == 2) m = ((e & 0xFFu) << 8) | (e >> 8);
  else if (BPE == 3) m = ((e & 0xFFu) << 16) | (e & 0xFF00u) | ((e >> 16) & 0xFFu);
  else m = __builtin_bswap32(e);
  const uint32_t sh = (off & 3u) * 8u;
  atomicOr(&s_win[off >> 2], m << sh);
  if (BPE == 3 && (off & 3u) > 1u) atomicOr(&s_win[(off >> 2) + 1], m >> (32u - sh));
}

// Copies table bytes [B0, B1) from the window, whose word 0 holds the table
// bytes from the 16-byte boundary at or below B0 (absolute addresses): whole
// 16-byte stores inside, byte stores in the two edge pieces.  skip_zero: the
// table is all zeros already, so zero pieces are not stored.  Each piece is
// cleared once copied, so the window is all zeros again afterwards.
__device__ __forceinline__ void win_out(uint8_t* __restrict__ t, uint32_t* s_win, uint64_t B0, uint64_t B1,
                                        bool skip_zero) {
  const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(t) + B0) & 15u);
  uint8_t* base = t + B0 - mis;  // 16-byte aligned; bytes before t + B0 are never stored
  const uint32_t nb = (uint32_t)(B1 - B0) + mis;
  const uint32_t nch = (nb + 15) / 16;
  for (uint32_t k = threadIdx.x; k < nch; k += RB_THREADS) {
    const uint4 x = reinterpret_cast<const uint4*>(s_win)[k];
    const uint32_t lo = 16 * k, hi = 16 * k + 16;
    if (lo >= mis && hi <= nb) {
      if (!skip_zero || (x.x | x.y | x.z | x.w)) *reinterpret_cast<uint4*>(base + lo) = x;
    } else {
      const uint8_t* xb = reinterpret_cast<const uint8_t*>(s_win) + lo;
      for (uint32_t i = lo < mis ? mis : lo; i < (hi < nb ? hi : nb); ++i) base[i] = xb[i - lo];
    }
    reinterpret_cast<uint4*>(s_win)[k] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// Clears a write window (each thread its 16-byte pieces); rb_emit expects a
// cleared window and leaves it cleared.
template <int BPE>
__device__ __forceinline__ void win_clear(uint32_t* s_win) {
  for (uint32_t k = threadIdx.x; k < rb_win_words<BPE>() / 4; k += RB_THREADS)
    reinterpret_cast<uint4*>(s_win)[k] = make_uint4(0u, 0u, 0u, 0u);
}

// Writes one sub-tile: block scan of the per-thread states (v, keepbits, a)
// on top of the incoming state X; the sub-tile owns the slots [X.lp, Y.lp)
// up to its last kept entry and writes every byte of them (elements and the
// zeros between them) window by window: every kept entry ORs its element
// into the LDS window (all zeros on entry) and every run head its occupied
// bit (ORs commute, so no ordering among threads), then the window is copied
// out and cleared again.  Run heads whose quotient precedes the owned range go to the fix-up
// list.  precleared: the table is all zeros, so windows without an element
// or occupied bit are skipped.  Every thread of the block calls it; returns
// the outgoing state.
template <int BPE>
__device__ __forceinline__ RS rb_emit(uint8_t* __restrict__ t, uint32_t q, uint32_t r, const uint32_t* v,
                                      uint32_t keepbits, RS a, RS X, uint32_t u, uint4* s_wave,
                                      uint32_t* s_win, uint32_t* s_nfix_p,
                                      uint32_t* __restrict__ fix, uint32_t* __restrict__ fix_cnt,
                                      uint32_t* __restrict__ fix_more, uint32_t* __restrict__ fix_more_n,
                                      uint32_t fix_cap, int precleared) {
  const int tid = threadIdx.x;
  const uint32_t mq = (uint32_t)mask64(q);
  const uint32_t rmask = (1u << r) - 1u;
  constexpr uint32_t bpe = BPE;
  uint32_t& s_nfix = *s_nfix_p;
  if (tid == 0) s_nfix = 0;
  RS ex;
  const RS sub = rs_block(a, &ex, s_wave);
  const RS me = rs_combine(X, ex);  // state before this thread's entries
  const RS Y = rs_combine(X, sub);
  const uint32_t Llo = X.lp, Lhi = Y.lp;
  // per-thread location walk: prev = location of the previous kept entry
  // (-1: none), lastf = its quotient (0xFFFFFFFF: none)
  const int32_t prev0 = (int32_t)me.lp - 1;
  const uint32_t lastf0 = me.fq1 - 1u;
  constexpr uint32_t WS = rb_win_slots<BPE>();
  for (uint32_t w0 = Llo; w0 < Lhi; w0 += WS) {
    const uint32_t w1 = w0 + WS < Lhi ? w0 + WS : Lhi;
    if (precleared) {
      bool any = false;
      int32_t prev = prev0;
      uint32_t lastf = lastf0;
#pragma unroll
      for (int j = 0; j < RB_ITEMS; ++j) {
        if (keepbits & (1u << j)) {
          const uint32_t fq = (v[1 + j] >> r) & mq;
          const int32_t lc = (int32_t)fq > prev + 1 ? (int32_t)fq : prev + 1;
          any |= ((uint32_t)lc - w0 < w1 - w0) || (fq != lastf && fq - w0 < w1 - w0);
          prev = lc;
          lastf = fq;
        }
      }
      if (!__syncthreads_or(any)) continue;
    }
    const uint64_t B0 = (uint64_t)w0 * bpe;
    const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(t) + B0) & 15u);
    {
      // window byte of slot s: (s - w0) * bpe + mis
      int32_t prev = prev0;
      uint32_t lastf = lastf0;
#pragma unroll
      for (int j = 0; j < RB_ITEMS; ++j) {
        const uint32_t fp = v[1 + j];
        const uint32_t fq = (fp >> r) & mq;
        const bool kp = (keepbits >> j) & 1u;
        const int32_t lc = (int32_t)fq > prev + 1 ? (int32_t)fq : prev + 1;
        const bool head = fq != lastf;
        const bool in_e = kp && (uint32_t)lc - w0 < w1 - w0;
        const bool in_o = kp && head && fq - w0 < w1 - w0;
        const uint32_t e = ((fp & rmask) << 3) | (head ? 0u : 2u) | ((uint32_t)lc != fq ? 1u : 0u);
        const uint32_t eoff = ((uint32_t)lc - w0) * bpe + mis;
        if (in_e) win_or<BPE>(s_win, eoff, e);
        if (in_o) {
          const uint32_t o = (fq - w0) * bpe + mis + bpe - 1;
          atomicOr(&s_win[o >> 2], 4u << ((o & 3u) * 8u));
        }
        if (in_e && head && fq < Llo) {
          // the run head's quotient precedes the owned range
          const uint32_t k = atomicAdd(&s_nfix, 1u);
          if (k < RB_FIX) {
            fix[(uint64_t)u * RB_FIX + k] = fq;
          } else {
            const uint32_t slot = atomicAdd(fix_more_n, 1u);
            if (slot < fix_cap) fix_more[slot] = fq;
          }
        }
        if (kp) {
          prev = lc;
          lastf = fq;
        }
      }
    }
    __syncthreads();
    win_out(t, s_win, B0, (uint64_t)w1 * bpe, precleared != 0);
    __syncthreads();
  }
  if (tid == 0) fix_cnt[u] = s_nfix < RB_FIX ? s_nfix : RB_FIX;
  return Y;
}

template <int MODE, bool MERGE, int BPE>
__global__ __launch_bounds__(RB_THREADS, 6) void k_rb_write(uint8_t* __restrict__ t, uint32_t S, uint32_t q, uint32_t r,
                                                  const uint32_t* __restrict__ A, uint32_t na,
                                                  const uint32_t* __restrict__ na_dev,
                                                  const uint32_t* __restrict__ B, uint32_t nb,
                                                  const uint8_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ splits, uint32_t chunk,
                                                  const uint4* __restrict__ pre,
                                                  const uint32_t* __restrict__ ctr, uint32_t* __restrict__ fix,
                                                  uint32_t* __restrict__ fix_cnt, uint32_t* __restrict__ fix_more,
                                                  uint32_t* __restrict__ fix_more_n, uint32_t fix_cap,
                                                  int precleared) {
  if (ctr[CTR_ERR]) return;  // capacity error: the table stays as it was
  __shared__ uint32_t s_stage[MERGE ? 2 : 1][MERGE ? RB_TILE : 1];
  __shared__ uint32_t s_prev[2][2];
  __shared__ uint32_t s_bmp[MERGE ? RB_TILE / 32 : 1], s_wpre[MERGE ? RB_TILE / 32 : 1];
  __shared__ __align__(16) uint32_t s_win[rb_win_words<BPE>()];
  __shared__ uint4 s_wave[RB_WAVES + 1];
  __shared__ uint32_t s_nfix;
  win_clear<BPE>(s_win);  // (the barriers before the first scatter order it)
  if (na_dev) na = *na_dev;
  const uint64_t N = (uint64_t)na + (MERGE ? nb : 0u);
  if (precleared < 0) precleared = N * 8 < (uint64_t)S ? 1 : 0;  // decided from the count on the device
  const uint64_t mq = mask64(q);
  const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t c1 = c0 + chunk < N ? c0 + chunk : N;
  RS X = rs_unpack(pre[blockIdx.x]);
  // splits: ca* of the current sub-tile, na* of the next one (loaded a
  // sub-tile ahead, so that issuing its stage does not wait on them)
  uint32_t ca0 = 0, ca1 = 0, na0 = 0, na1 = 0;
  if (MERGE && c0 < c1) {
    const uint32_t u0 = (uint32_t)(c0 / RB_TILE);
    ca0 = splits[u0];
    ca1 = splits[u0 + 1];
    if (c0 + RB_TILE < c1) na1 = splits[u0 + 2];
    na0 = ca1;
    rb_stage(A, B, ca0, ca1, (uint32_t)c0, (uint32_t)(c1 - c0 < RB_TILE ? c1 - c0 : RB_TILE), s_stage[0],
                   s_prev[0]);
  }
  int buf = 0;
  for (uint64_t d0 = c0; d0 < c1; d0 += RB_TILE) {
    const uint32_t u = (uint32_t)(d0 / RB_TILE);
    const uint32_t tn = (uint32_t)(c1 - d0 < (uint64_t)RB_TILE ? c1 - d0 : (uint64_t)RB_TILE);
    uint32_t v[RB_ITEMS + 1];
    uint32_t tags, fl = 0;
    if (MERGE) {
      const uint64_t dn = d0 + RB_TILE;
      rb_stage_wait();
      // the previous merge's bitmap reads ended before the barriers since
      if (threadIdx.x < RB_TILE / 32) s_bmp[threadIdx.x] = 0u;
      __syncthreads();
      if (dn < c1)
        rb_stage(A, B, na0, na1, (uint32_t)dn, (uint32_t)(c1 - dn < RB_TILE ? c1 - dn : RB_TILE),
                       s_stage[buf ^ 1], s_prev[buf ^ 1]);
      const uint32_t nn1 = dn + RB_TILE < c1 ? splits[u + 3] : 0u;
      if (tn == RB_TILE)
        rb_merge_staged<true>(s_stage[buf], s_prev[buf], ca0, ca1, (uint32_t)d0, tn, v, tags, s_bmp, s_wpre);
      else
        rb_merge_staged<false>(s_stage[buf], s_prev[buf], ca0, ca1, (uint32_t)d0, tn, v, tags, s_bmp, s_wpre);
      ca0 = na0;
      ca1 = na1;
      na0 = na1;
      na1 = nn1;
      buf ^= 1;
    } else {
      rb_load<MODE>(A, flags, (uint32_t)d0, tn, v, tags, fl);
    }
    RS a;
    const uint32_t keepbits = tn == RB_TILE ? rb_classify<MODE, true>(v, tags, fl, (uint32_t)d0, tn, r, mq, a)
                                            : rb_classify<MODE, false>(v, tags, fl, (uint32_t)d0, tn, r, mq, a);
    X = rb_emit<BPE>(t, q, r, v, keepbits, a, X, u, s_wave, s_win, &s_nfix, fix, fix_cnt, fix_more,
                     fix_more_n, fix_cap, precleared);
    __syncthreads();
  }
}
// Zeroes every table byte from the slot after the last entry to the end of
// the allocation (the 10% overflow area, usually), in parallel.
// (whole_table: every byte, used to pre-clear sparse tables).  Nothing is
// cleared after a capacity error.
// mode 0: the tail; 1: the whole table; 2: the whole table iff the rebuild
// is sparse (N * 8 < S, then the writer skips windows of zeros); 3: the tail
// iff it is not.  N = *n_dev (when given) + n_add.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_sqf_clear_tail(uint8_t* __restrict__ t, uint32_t bpe, uint64_t tbytes,
                                                        const uint32_t* __restrict__ ctr, int mode,
                                                        const uint32_t* __restrict__ n_dev, uint32_t n_add,
                                                        uint32_t S) {
  if (ctr[CTR_ERR]) return;  // capacity error: the table stays as it was
  if (mode >= 2) {
    const uint64_t N = (uint64_t)(n_dev ? *n_dev : 0u) + n_add;
    const bool sparse = N * 8 < (uint64_t)S;
    if (sparse != (mode == 2)) return;
  }
  const int whole_table = mode == 1 || mode == 2;
  const uint64_t b0 = whole_table ? 0ull : (uint64_t)ctr[CTR_MISC + 3] * bpe;
  const uint64_t a0 = (b0 + 15) & ~15ull, a1 = tbytes & ~15ull;
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x, nth = (uint64_t)gridDim.x * 256;
  if (a0 >= a1) {
    for (uint64_t b = b0 + tid; b < tbytes; b += nth) t[b] = 0;
    return;
  }
  for (uint64_t b = b0 + tid; b < a0; b += nth) t[b] = 0;
  for (uint64_t b = a1 + tid; b < tbytes; b += nth) t[b] = 0;
  uint4* w = reinterpret_cast<uint4*>(t + a0);
  const uint64_t nw = (a1 - a0) / 16;
  for (uint64_t i = tid; i < nw; i += nth) w[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Occupied bits that rebuild tiles could not place in their own ranges: one
// workgroup per tile's slots, plus the shared overflow list.
__device__ __forceinline__ void sqf_set_occupied(uint8_t* t, uint32_t bpe, uint32_t fq) {
  const uint64_t byte = (uint64_t)fq * bpe + bpe - 1;  // the element's low byte holds the flag bits
  uint32_t* word = reinterpret_cast<uint32_t*>(t + (byte & ~3ull));
  atomicOr(word, 4u << (8 * (uint32_t)(byte & 3)));
}

[[maybe_unused]] static __global__ __launch_bounds__(64) void k_sqf_fixup(uint8_t* __restrict__ t, uint32_t bpe,
                                                  const uint32_t* __restrict__ fix,
                                                  const uint32_t* __restrict__ fix_cnt,
                                                  const uint32_t* __restrict__ fix_more,
                                                  const uint32_t* __restrict__ ctr, uint32_t tiles,
                                                  uint32_t fix_cap, const uint32_t* __restrict__ n_dev,
                                                  uint32_t n_add) {
  if (ctr[CTR_ERR]) return;  // capacity error: the table stays as it was
  const uint32_t tile = blockIdx.x;
  if (n_dev) {  // the grid covers an upper bound: sub-tiles past the entries have no list
    const uint64_t N = (uint64_t)*n_dev + n_add;
    const uint64_t used = (N + RB_TILE - 1) / RB_TILE;
    if (tile < tiles && tile >= used) return;
  }
  if (tile < tiles) {
    const uint32_t c = fix_cnt[tile];
    for (uint32_t i = threadIdx.x; i < c; i += 64) sqf_set_occupied(t, bpe, fix[(uint64_t)tile * RB_FIX + i]);
  } else {
    uint32_t n = ctr[CTR_MISC + 2];  // shared overflow list length
    if (n > fix_cap) n = fix_cap;
    for (uint32_t i = threadIdx.x; i < n; i += 64) sqf_set_occupied(t, bpe, fix_more[i]);
  }
}

// ---------------------------------------------------------------- decode
// Every stored fingerprint in slot (= ascending) order.  The k-th run start of
// the table belongs to the k-th occupied quotient, so a region decodes from
// three exclusive counts (entries, run starts, occupied bits) before it, plus
// the occupied positions just before it that still own runs.  Replaces
// extractQuotients + remove_if (quotientFilter.cu:1561-1600, 1627).
constexpr uint32_t DC_SLOTS = 4096;
// Pending quotients kept in LDS (the rest are found by walking the cluster);
// 512 keeps k_sqf_decode within 6 workgroups per CU.
constexpr uint32_t DC_PEND = 512;

// The 16 slots [s, s+16) of a thread: 16-byte loads when whole, else
// bounded byte loads (past the last slot reads as empty).
template <int BPE>
__device__ __forceinline__ void sqf_load16(const uint8_t* __restrict__ t, uint32_t S, uint32_t s, uint32_t* e) {
  if (s + 16 <= S && (reinterpret_cast<uintptr_t>(t) & 15u) == 0) {
    uint32_t w[4 * BPE];
    const uint4* p = reinterpret_cast<const uint4*>(t + (uint64_t)s * BPE);
#pragma unroll
    for (int k = 0; k < BPE; ++k) {
      const uint4 x = p[k];
      w[4 * k] = x.x;
      w[4 * k + 1] = x.y;
      w[4 * k + 2] = x.z;
      w[4 * k + 3] = x.w;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < BPE; ++b) {
        const int at = j * BPE + b;
        v = (v << 8) | ((w[at >> 2] >> (8 * (at & 3))) & 0xFFu);
      }
      e[j] = v;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) e[j] = sqf_get<BPE>(t, S, s + j);
  }
}

// Entries, run starts and occupieds among the 4 one-byte slots of a table
// word (byte b = slot 4k + b), all at once: per byte, bit 2 = occupied,
// bit 1 = continuation, any of bits 0..2 = holds an entry.
__device__ __forceinline__ void sqf_flags4(uint32_t w, uint32_t& ci, uint32_t& cr, uint32_t& co) {
  const uint32_t f = w & 0x07070707u;
  const uint32_t nz = (f | (f >> 1) | (f >> 2)) & 0x01010101u;
  ci += (uint32_t)__popc(nz);
  cr += (uint32_t)__popc(nz & ~(w >> 1));
  co += (uint32_t)__popc(w & 0x04040404u);
}

// The 16 one-byte slots [s, s+16) as 4 words when whole and aligned.
__device__ __forceinline__ bool sqf_words16(const uint8_t* __restrict__ t, uint32_t S, uint32_t s, uint32_t (&w)[4]) {
  if (s + 16 > S || (reinterpret_cast<uintptr_t>(t) & 15u) != 0) return false;
  const uint4 x = *reinterpret_cast<const uint4*>(t + s);
  w[0] = x.x;
  w[1] = x.y;
  w[2] = x.z;
  w[3] = x.w;
  return true;
}

template <int BPE>
__global__ __launch_bounds__(256) void k_sqf_count(const uint8_t* __restrict__ t, uint32_t S,
                                                   uint32_t* __restrict__ counts) {
  __shared__ uint64_t s_tmp[4];
  const uint32_t s0 = blockIdx.x * DC_SLOTS + threadIdx.x * 16;
  uint32_t ci = 0, cr = 0, co = 0;
  uint32_t w[4];
  if (BPE == 1 && sqf_words16(t, S, s0, w)) {
#pragma unroll
    for (int k = 0; k < 4; ++k) sqf_flags4(w[k], ci, cr, co);
  } else {
    uint32_t e[16];
    sqf_load16<BPE>(t, S, s0, e);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      ci += (e[j] & 7u) != 0;
      cr += ((e[j] & 7u) != 0) && !(e[j] & 2u);
      co += (e[j] & 4u) != 0;
    }
  }
  uint64_t tot;
  block256_excl_scan64((uint64_t)ci | ((uint64_t)cr << 21) | ((uint64_t)co << 42), s_tmp, &tot);
  if (threadIdx.x == 0) {
    counts[blockIdx.x * 3 + 0] = (uint32_t)(tot & 0x1FFFFF);
    counts[blockIdx.x * 3 + 1] = (uint32_t)((tot >> 21) & 0x1FFFFF);
    counts[blockIdx.x * 3 + 2] = (uint32_t)(tot >> 42);
  }
}

// Quotient of the entry at slot idx by walking its cluster (extractQuotients,
// quotientFilter.cu:1580-1599); fallback for very long pending chains, kept
// out of line so that the decode's unrolled entry loop holds one call, not
// sixteen copies of the walk.
template <int BPE>
__device__ __noinline__ uint32_t sqf_walk_quotient(const uint8_t* __restrict__ t, uint32_t S, uint32_t idx) {
  uint32_t b = idx;
  do {
    b--;
  } while (b > 0 && (sqf_get<BPE>(t, S, b) & 1u));
  uint32_t s = b;
  while (s <= idx) {
    do {
      s++;
    } while (sqf_get<BPE>(t, S, s) & 2u);
    if (s > idx) break;
    do {
      b++;
    } while (!(sqf_get<BPE>(t, S, b) & 4u) && b < S);
  }
  return b;
}

template <int BPE, typename K>
__global__ __launch_bounds__(256) void k_sqf_decode(const uint8_t* __restrict__ t, uint32_t S, uint32_t r,
                                                    const uint32_t* __restrict__ bases, K* __restrict__ out) {
  __shared__ uint16_t s_occpos[DC_SLOTS];
  __shared__ uint32_t s_pend[DC_PEND];
  __shared__ K s_out[DC_SLOTS];
  __shared__ uint64_t s_tmp[4];
  __shared__ uint32_t s_tmp32[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t s0 = blockIdx.x * DC_SLOTS;
  const uint32_t bI = bases[blockIdx.x * 3 + 0], bR = bases[blockIdx.x * 3 + 1], bO = bases[blockIdx.x * 3 + 2];
  uint32_t e[16];
  uint32_t ci = 0, cr = 0, co = 0;
  uint32_t w[4];
  if (BPE == 1 && sqf_words16(t, S, s0 + tid * 16, w)) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sqf_flags4(w[k], ci, cr, co);
#pragma unroll
      for (int b = 0; b < 4; ++b) e[4 * k + b] = (w[k] >> (8 * b)) & 0xFFu;
    }
  } else {
    sqf_load16<BPE>(t, S, s0 + tid * 16, e);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      ci += (e[j] & 7u) != 0;
      cr += ((e[j] & 7u) != 0) && !(e[j] & 2u);
      co += (e[j] & 4u) != 0;
    }
  }
  uint64_t btot;
  const uint64_t ex = block256_excl_scan64((uint64_t)ci | ((uint64_t)cr << 21) | ((uint64_t)co << 42), s_tmp, &btot);
  const uint32_t nout = (uint32_t)(btot & 0x1FFFFF);
  uint32_t xi = (uint32_t)(ex & 0x1FFFFF), xr = (uint32_t)((ex >> 21) & 0x1FFFFF), xo = (uint32_t)(ex >> 42);
  {
    uint32_t o = xo;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (e[j] & 4u) s_occpos[o++] = (uint16_t)(tid * 16 + j);
  }
  // occupied slots before s0 whose runs start at or after s0 (nearest first):
  // bO - bR of them, plus the owner of a run that continues into the region
  __shared__ uint32_t s_first;
  if (tid == 0) s_first = e[0];
  __syncthreads();
  const uint32_t want = (bO >= bR ? bO - bR : 0u) + ((s_first & 2u) ? 1u : 0u);
  uint32_t got = 0;
  for (uint32_t c = s0; got < want && got < DC_PEND && c > 0;) {
    const uint32_t lo = c > 256 ? c - 256 : 0u;
    const uint32_t slot = lo + tid;
    const uint32_t isocc = (slot < c) && (sqf_get<BPE>(t, S, slot) & 4u) ? 1u : 0u;
    uint32_t total;
    const uint32_t below = block256_excl_scan(isocc, s_tmp32, &total);
    if (isocc) {
      const uint32_t fromEnd = got + (total - below - 1);
      if (fromEnd < DC_PEND) s_pend[fromEnd] = slot;
    }
    got += total;
    c = lo;
  }
  if (got > DC_PEND) got = DC_PEND;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint32_t el = e[j];
    if (!(el & 7u)) continue;
    if (!(el & 2u)) xr++;
    const uint32_t k = bR + xr;  // 1-based run index of this entry
    uint32_t fq;
    if (k > bO) {
      fq = s0 + s_occpos[k - bO - 1];
    } else if (k >= 1 && bO - k < got) {
      fq = s_pend[bO - k];
    } else {
      fq = sqf_walk_quotient<BPE>(t, S, s0 + tid * 16 + j);
    }
    if constexpr (sizeof(K) == 4) s_out[xi] = (fq << r) | (el >> 3);  // q + r <= 32
    else s_out[xi] = (K)(((uint64_t)fq << r) | (el >> 3));
    xi++;
  }
  __syncthreads();
  for (uint32_t k = tid; k < nout; k += 256) out[bI + k] = s_out[k];
}

// ---------------------------------------------------------------- localized insert
// A batch that is small next to the table is inserted by rewriting only the
// regions it changes (insert keep rule only).  In a canonical table the
// occupied slots after inserting a sorted batch are what linear probing
// gives: key i takes the first empty slot at or after max(h_i, p_{i-1} + 1)
// (h_i: its quotient).  With E(x) = the empty slots before x, that slot is
// the j_i-th empty slot, j_i = max(E(h_i), j_{i-1} + 1) -- the location
// recurrence over empty ranks, scanned with the same state as locations.  A
// boundary is an empty or unshifted slot (no run crosses it); key i's region
// starts at the last boundary at or before h_i, and a key whose boundary lies
// after the previous key's slot starts a new region, which ends at the last
// slot its keys take.  Slots before a region and after it keep their entries,
// so every region is decoded, merged with its keys and re-encoded in place,
// each by one thread.  A key's slot past the last empty slot, or a region
// longer than PI_LEN_CAP, sends the call to the full rebuild instead (flags).
// E(x) and the j-th empty slot come from per-region and per-64-slot-window
// prefixes plus one 64-slot bitmask, so each key costs a few loads.
constexpr uint32_t PI_REGION = 4096;        // slots per counting region (64 windows of 64)
constexpr uint32_t PI_LEN_CAP = 1u << 16;   // longest region a thread rewrites
constexpr uint32_t PI_SCAN_CAP = 1u << 16;  // longest backward scan to a boundary
// Automatic choice: batches of at most PI_AUTO_MAX keys and S / 256 of the
// table.  Measured on a q=26 r=5 table (ms per call, localized / rebuild):
// at fill 0.45, 2^14 keys 0.43 / 0.57, 2^16 0.43 / 0.55, 2^18 0.50 / 0.55;
// at fill 0.75, 2^16 0.64 / 0.63, 2^17 0.68 / 0.62 -- its regions grow with
// the fill, and one thread rewrites each.
constexpr uint32_t PI_AUTO_MAX = 65536;
constexpr uint32_t PI_MIN_RATIO = 256;

template <int BPE>
__device__ __forceinline__ void sqf_put(uint8_t* __restrict__ t, uint32_t i, uint32_t e) {
  uint8_t* p = t + (size_t)i * BPE;
#pragma unroll
  for (int b = 0; b < BPE; ++b) p[b] = (uint8_t)(e >> (8 * (BPE - 1 - b)));
}

// Empty slots per region (rcnt) and, per 64-slot window, the empty slots of
// its region before it (wpre, 16 bits).
template <int BPE>
__global__ __launch_bounds__(256) void k_pi_count(const uint8_t* __restrict__ t, uint32_t S, uint16_t* __restrict__ wpre,
                                                  uint32_t* __restrict__ rcnt) {
  __shared__ uint32_t s_tmp4[4];
  const uint32_t s0 = blockIdx.x * PI_REGION + threadIdx.x * 16;
  uint32_t e[16];
  sqf_load16<BPE>(t, S, s0, e);
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) c += (s0 + (uint32_t)j < S && (e[j] & 7u) == 0) ? 1u : 0u;
  uint32_t tot;
  const uint32_t pre = block256_excl_scan(c, s_tmp4, &tot);
  if ((threadIdx.x & 3) == 0 && s0 < S) wpre[s0 >> 6] = (uint16_t)pre;  // the window's first thread
  if (threadIdx.x == 0) rcnt[blockIdx.x] = tot;
}

// Bit j set: slot 64 w + j is empty (*empty) / a boundary -- empty or
// unshifted, so that no run crosses it (*bound).  Slots past S: neither.
template <int BPE>
__device__ __forceinline__ void pi_masks64(const uint8_t* __restrict__ t, uint32_t S, uint32_t w, uint64_t* empty,
                                           uint64_t* bound) {
  uint64_t em = 0, bd = 0;
#pragma unroll
  for (uint32_t c = 0; c < 4; ++c) {
    uint32_t e[16];
    sqf_load16<BPE>(t, S, 64 * w + 16 * c, e);
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
      if (64 * w + 16 * c + j < S) {
        em |= (uint64_t)((e[j] & 7u) == 0) << (16 * c + j);
        bd |= (uint64_t)((e[j] & 1u) == 0) << (16 * c + j);
      }
    }
  }
  *empty = em;
  *bound = bd;
}

// Empty slots before window w (ER: exclusive region prefix of rcnt).
__device__ __forceinline__ uint32_t pi_empties_before_window(const uint16_t* __restrict__ wpre,
                                                             const uint32_t* __restrict__ ER, uint32_t w) {
  return ER[w / (PI_REGION / 64)] + wpre[w];
}

// The j-th empty slot (0-based), known to lie at or after slot h: windows
// are walked forward from h's (usually it is there or in the next one).
template <int BPE>
__device__ uint32_t pi_select_empty(const uint8_t* __restrict__ t, uint32_t S, const uint16_t* __restrict__ wpre,
                                    const uint32_t* __restrict__ ER, uint32_t h, uint32_t j) {
  const uint32_t wend = (S + 63) / 64;
  uint32_t w = h >> 6;
  if (w >= wend) return S;
  uint32_t g = pi_empties_before_window(wpre, ER, w);
  for (; w < wend; ++w) {
    uint64_t em, bd;
    pi_masks64<BPE>(t, S, w, &em, &bd);
    const uint32_t c = (uint32_t)__popcll(em);
    if (g + c > j) return 64 * w + select_bit64(em, j - g + 1);
    g += c;
  }
  return S;  // not reached for a consistent count
}

// Per key: empties before its quotient, and its region's boundary (the last
// boundary slot at or before it).
template <int BPE>
__global__ __launch_bounds__(256) void k_pi_keys(const uint8_t* __restrict__ t, uint32_t S, uint32_t q, uint32_t r,
                                                 const uint32_t* __restrict__ B, uint32_t nb,
                                                 const uint16_t* __restrict__ wpre, const uint32_t* __restrict__ ER,
                                                 uint32_t* __restrict__ E, uint32_t* __restrict__ s0,
                                                 uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nb) return;
  const uint32_t h = (B[i] >> r) & (uint32_t)mask64(q);
  uint32_t w = h >> 6;
  uint64_t em, bd;
  pi_masks64<BPE>(t, S, w, &em, &bd);
  const uint64_t below = (1ull << (h & 63)) - 1ull;
  E[i] = pi_empties_before_window(wpre, ER, w) + (uint32_t)__popcll(em & below);
  uint64_t m = bd & (below | (1ull << (h & 63)));  // boundaries at or before h
  uint32_t steps = 0;
  while (!m && w > 0) {
    --w;
    pi_masks64<BPE>(t, S, w, &em, &bd);
    m = bd;
    if (++steps * 64 > PI_SCAN_CAP) {
      atomicOr(flags, 1u);
      break;
    }
  }
  s0[i] = m ? 64 * w + 63 - (uint32_t)__clzll(m) : 0u;
}

// Chunk states of the slot-rank recurrence (RB_TILE keys per workgroup, as
// the rebuild's location scan; the "quotient" of key i is E_i).
[[maybe_unused]] static __global__ __launch_bounds__(RB_THREADS) void k_pi_jagg(const uint32_t* __restrict__ E, uint32_t nb,
                                                        uint4* __restrict__ agg) {
  __shared__ uint4 s_wave[RB_WAVES + 1];
  const uint32_t i0 = blockIdx.x * RB_TILE + threadIdx.x * RB_ITEMS;
  RS a = rs_none();
#pragma unroll
  for (int k = 0; k < RB_ITEMS; ++k)
    if (i0 + k < nb) rs_push(a, E[i0 + k]);
  RS ex;
  const RS sub = rs_block(a, &ex, s_wave);
  if (threadIdx.x == 0) agg[blockIdx.x] = rs_pack(sub);
}

// Each key's empty rank j_i = max(E_i, j_{i-1} + 1), from the chunk prefixes
// of the rank scan (RB_TILE keys per workgroup).
[[maybe_unused]] static __global__ __launch_bounds__(RB_THREADS) void k_pi_ranks(const uint32_t* __restrict__ E,
                                                                                 uint32_t nb,
                                                                                 const uint4* __restrict__ pre,
                                                                                 const uint32_t* __restrict__ ER,
                                                                                 uint32_t R, uint32_t* __restrict__ J,
                                                                                 uint32_t* __restrict__ flags) {
  __shared__ uint4 s_wave[RB_WAVES + 1];
  const uint32_t tid = threadIdx.x, i0 = blockIdx.x * RB_TILE + tid * RB_ITEMS;
  const uint32_t total = ER[R];
  uint32_t e[RB_ITEMS];
  RS a = rs_none();
#pragma unroll
  for (int k = 0; k < RB_ITEMS; ++k) {
    e[k] = i0 + k < nb ? E[i0 + k] : 0u;
    if (i0 + k < nb) rs_push(a, e[k]);
  }
  RS ex;
  (void)rs_block(a, &ex, s_wave);
  uint32_t lp = rs_combine(rs_unpack(pre[blockIdx.x]), ex).lp;
#pragma unroll
  for (int k = 0; k < RB_ITEMS; ++k) {
    if (i0 + k < nb) {
      lp = lp + 1u > e[k] + 1u ? lp + 1u : e[k] + 1u;
      if (lp - 1u >= total) atomicOr(flags, 2u);  // no empty slot left: the full rebuild decides
      J[i0 + k] = lp - 1u;
    }
  }
}

// One thread per key: its slot p_i (the j_i-th empty slot, at or after its
// quotient) and whether it heads a region (i == 0 or s0_i > p_{i-1}; the
// previous key's slot is selected again here rather than waited for).
template <int BPE>
__global__ __launch_bounds__(256) void k_pi_pos(const uint8_t* __restrict__ t, uint32_t S, uint32_t q, uint32_t r,
                                                const uint32_t* __restrict__ B, const uint32_t* __restrict__ J,
                                                const uint32_t* __restrict__ s0, uint32_t nb,
                                                const uint16_t* __restrict__ wpre, const uint32_t* __restrict__ ER,
                                                uint32_t R, uint32_t* __restrict__ P, uint8_t* __restrict__ head) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nb) return;
  const uint32_t total = ER[R], mq = (uint32_t)mask64(q);
  const uint32_t j = J[i];
  const uint32_t p = j < total ? pi_select_empty<BPE>(t, S, wpre, ER, (B[i] >> r) & mq, j) : S;
  P[i] = p;
  bool h = true;
  if (i > 0) {
    const uint32_t jp = J[i - 1];
    const uint32_t pp = jp < total ? pi_select_empty<BPE>(t, S, wpre, ER, (B[i - 1] >> r) & mq, jp) : S;
    h = s0[i] > pp;
  }
  head[i] = h ? 1 : 0;
}

// Region of each head key: its keys [i, rend_i) and a length check.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_pi_regions(const uint32_t* __restrict__ s0, const uint32_t* __restrict__ P,
                                                    const uint8_t* __restrict__ head, uint32_t nb,
                                                    uint32_t* __restrict__ rend, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nb || !head[i]) return;
  uint32_t k = i + 1;
  while (k < nb && !head[k] && k - i < PI_LEN_CAP) ++k;
  rend[i] = k;
  if ((k < nb && !head[k]) || P[k - 1] - s0[i] + 1u > PI_LEN_CAP) atomicOr(flags, 4u);
}

[[maybe_unused]] static __global__ void k_pi_init(uint32_t* __restrict__ rcnt_last, uint32_t* __restrict__ flags) {
  if (threadIdx.x == 0) {
    *rcnt_last = 0u;
    *flags = 0u;
  }
}

// One thread per region [s0_i, P[rend_i - 1]]: decode its entries into the
// slot-indexed scratch, merge the region's keys (insert rule: a key equal
// to the entry before it is dropped), write every slot of the region (elements
// and zeros), then the occupied bits of its run heads.  (One workgroup per
// 4096-slot window rewriting its regions from an LDS copy was measured
// slower: the windows run in rounds, each as long as its longest region,
// where one thread per key keeps every region in flight at once.)
template <int BPE>
__global__ __launch_bounds__(256) void k_pi_write(uint8_t* __restrict__ t, uint32_t S, uint32_t r,
                                                  const uint32_t* __restrict__ B, uint32_t nb,
                                                  const uint32_t* __restrict__ s0, const uint32_t* __restrict__ P,
                                                  const uint8_t* __restrict__ head, const uint32_t* __restrict__ rend,
                                                  uint32_t* __restrict__ scratch, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nb || !head[i]) return;
  const uint32_t lo = s0[i], kend = rend[i], hi = P[kend - 1];
  const uint32_t rmask = (1u << r) - 1u;
  // decode: an unshifted entry starts a run at its own slot; a shifted run
  // start takes the next occupied quotient after the current one
  uint32_t m = 0, qcur = 0;
  for (uint32_t x = lo; x <= hi; ++x) {
    const uint32_t ex = sqf_get<BPE>(t, S, x);
    if ((ex & 7u) == 0) continue;
    if (!(ex & 1u)) {
      qcur = x;
    } else if (!(ex & 2u)) {
      uint32_t y = qcur + 1;
      while (y <= hi && !(sqf_get<BPE>(t, S, y) & 4u)) ++y;
      if (y > hi) {
        atomicOr(flags, 8u);  // inconsistent table: leave the region as it was
        return;
      }
      qcur = y;
    }
    scratch[lo + m++] = (qcur << r) | (ex >> 3);
  }
  // merge + encode; pass 0 writes elements and zeros, pass 1 the occupied bits
  for (int pass = 0; pass < 2; ++pass) {
    uint32_t ia = 0, ib = i, next = lo, lastq = 0xFFFFFFFFu;
    int64_t loc = (int64_t)lo - 1;
    uint64_t prevfp = ~0ull;
    while (ia < m || ib < kend) {
      const bool takeA = ib >= kend || (ia < m && scratch[lo + ia] <= B[ib]);
      const uint32_t v = takeA ? scratch[lo + ia++] : B[ib++];
      const bool keep = takeA || prevfp != (uint64_t)v;
      prevfp = v;
      if (!keep) continue;
      const uint32_t fq = v >> r;
      loc = (int64_t)fq > loc + 1 ? (int64_t)fq : loc + 1;
      if (loc > (int64_t)hi) {
        atomicOr(flags, 8u);  // not reached: the region holds its layout
        break;
      }
      if (pass == 0) {
        for (; next < (uint32_t)loc; ++next) sqf_put<BPE>(t, next, 0u);
        sqf_put<BPE>(t, (uint32_t)loc, ((v & rmask) << 3) | (fq == lastq ? 2u : 0u) | ((uint32_t)loc != fq ? 1u : 0u));
        next = (uint32_t)loc + 1;
      } else if (fq != lastq) {
        uint8_t* b = t + (size_t)fq * BPE + (BPE - 1);  // the element's low byte holds the flags
        *b = (uint8_t)(*b | 4u);
      }
      lastq = fq;
    }
    if (pass == 0)
      for (; next <= hi; ++next) sqf_put<BPE>(t, next, 0u);
  }
}

}  // namespace gqf
