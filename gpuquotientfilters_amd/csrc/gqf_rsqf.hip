// gqf_rsqf.hip -- rank-and-select quotient filter (RSQF) engine for gfx950.
//
// Device layout (one allocation, 256-byte aligned sections):
//   meta[NB]  32 bytes per 64-slot block: u64 occupieds, u64 runEnds,
//             u64 offset (low 8 bits; the reference stores a uint8), u64 0
//   rem[NB]   8*r bytes per block: slot s at bits [s*r, s*r+r), LSB first
//             (the bit order of the reference's remainders[] words,
//             RSQF.cu:91-125)
// so a lookup touches one 32-byte metadata record and the remainder bytes.
// gqf_rsqf_dump/load convert to and from the reference packed block
// (RSQF.cuh:24-30).
//
// Lookups restate lookupGPU (RSQF.cu:320-399) exactly, so they answer like the
// reference on any table (including the history-dependent layouts its insert
// leaves).  Build / insert / delete produce the canonical layout, which is
// what the reference insert produces when fed fingerprints in ascending order.
#include "gqf_rebuild.hpp"

#include <type_traits>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

namespace gqf {

constexpr uint32_t RS_NOT_FOUND = 0xFFFFFFFFu;

static inline uint32_t rsqf_blocks(unsigned q) { return (uint32_t)((((uint64_t)1 << q) + 63) / 64 + 1); }
__host__ __device__ static inline size_t rsqf_meta_bytes(uint32_t nb) { return aligned_bytes((size_t)nb * 32); }
static inline size_t rsqf_dev_bytes(unsigned q, unsigned r) {
  const uint32_t nb = rsqf_blocks(q);
  return rsqf_meta_bytes(nb) + (size_t)nb * 8 * r;
}
static inline size_t rsqf_ref_block_bytes(unsigned r) { return 1 + 8 + 8 + 8 * (size_t)r; }

static gqf_status rsqf_check(const gqf_rsqf& f) {
  if (!f.blocks) {
    set_error("RSQF blocks pointer is NULL (call gqf_rsqf_init)");
    return GQF_ERR_INVALID;
  }
  if (f.qbits < 6 || f.qbits > 30 || f.rbits < 1 || f.rbits > 16 || f.numBlocks != rsqf_blocks(f.qbits)) {
    set_error("unsupported RSQF geometry q=%u r=%u numBlocks=%u (need 6<=q<=30, 1<=r<=16)", f.qbits, f.rbits,
              f.numBlocks);
    return GQF_ERR_INVALID;
  }
  return GQF_OK;
}

// ---------------------------------------------------------------- block access
struct RsqfDev {
  const uint64_t* __restrict__ meta;
  const uint8_t* __restrict__ rem;
  uint32_t nb, r;
  __device__ __forceinline__ uint64_t occ(uint32_t b) const { return meta[4 * (size_t)b]; }
  __device__ __forceinline__ uint64_t ends(uint32_t b) const { return meta[4 * (size_t)b + 1]; }
  __device__ __forceinline__ uint32_t offset(uint32_t b) const { return (uint32_t)meta[4 * (size_t)b + 2] & 0xFFu; }
  // getRemainderGPU (RSQF.cu:91-107)
  template <int R>
  __device__ __forceinline__ uint32_t get(uint32_t b, uint32_t s) const {
    const uint32_t rr = R ? (uint32_t)R : r;
    if (rr == 8) return rem[(size_t)b * 64 + s];
    const uint8_t* p = rem + (size_t)b * 8 * rr;
    const uint32_t bit = s * rr;
    const uint32_t first = bit >> 3, last = (bit + rr - 1) >> 3;
    uint32_t v = p[first];
    if (last > first) v |= (uint32_t)p[first + 1] << 8;
    if (last > first + 1) v |= (uint32_t)p[first + 2] << 16;
    return (v >> (bit & 7)) & ((1u << rr) - 1u);
  }
};

// selectBitGPU (RSQF.cu:296-318) semantics: k == 0 -> 0; position of the k-th
// set bit; UINT_MAX if there are fewer.  Computed by halving with popcounts.
__device__ __forceinline__ uint32_t select_bit(uint64_t x, uint32_t k) {
  if (k == 0) return 0;
  if (k > (uint32_t)__popcll(x)) return RS_NOT_FOUND;
  uint32_t pos = 0, c;
  c = (uint32_t)__popc((uint32_t)x);
  if (k > c) { k -= c; x >>= 32; pos += 32; }
  c = (uint32_t)__popc((uint32_t)x & 0xFFFFu);
  if (k > c) { k -= c; x >>= 16; pos += 16; }
  c = (uint32_t)__popc((uint32_t)x & 0xFFu);
  if (k > c) { k -= c; x >>= 8; pos += 8; }
  c = (uint32_t)__popc((uint32_t)x & 0xFu);
  if (k > c) { k -= c; x >>= 4; pos += 4; }
  c = (uint32_t)__popc((uint32_t)x & 0x3u);
  if (k > c) { k -= c; x >>= 2; pos += 2; }
  c = (uint32_t)x & 1u;
  if (k > c) pos += 1;
  return pos;
}
__device__ __forceinline__ uint64_t low_mask_ll(uint32_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

// lookupGPU body (RSQF.cu:333-398), bounds-checked: blocks >= numBlocks end
// the search (the reference's check is `> numBlocks`).
template <int R>
__device__ __forceinline__ uint32_t rsqf_lookup_one(const RsqfDev& f, uint32_t fq, uint32_t fr) {
  uint32_t blockNum = fq >> 6;
  const uint32_t slotNum = fq & 63u;
  const uint64_t occs = f.occ(blockNum);
  if (!((occs >> slotNum) & 1ull)) return RS_NOT_FOUND;
  uint32_t blockOffset = f.offset(blockNum);
  uint32_t rank = (uint32_t)__popcll(occs & low_mask_ll(slotNum + 1));
  uint64_t ends = f.ends(blockNum);
  uint32_t endOfRun = select_bit(ends & ~low_mask_ll(blockOffset), rank);
  while (endOfRun == RS_NOT_FOUND) {
    rank -= (uint32_t)__popcll(ends & ~low_mask_ll(blockOffset));
    blockOffset = blockOffset > 64 ? blockOffset - 64 : 0u;
    blockNum++;
    if (blockNum >= f.nb) return RS_NOT_FOUND;
    ends = f.ends(blockNum);
    endOfRun = select_bit(ends & ~low_mask_ll(blockOffset), rank);
  }
  uint32_t slot = endOfRun;
  uint32_t cur = f.get<R>(blockNum, slot);
  for (;;) {
    if (cur == fr) return blockNum * 64 + slot;
    if (cur < fr) return RS_NOT_FOUND;
    if (slot > 0) {
      slot--;
    } else {
      slot = 63;
      if (blockNum == 0) return RS_NOT_FOUND;
      blockNum--;
      ends = f.ends(blockNum);
    }
    cur = f.get<R>(blockNum, slot);
    if (((ends >> slot) & 1ull) || (blockNum * 64 + slot) < fq) break;
  }
  return RS_NOT_FOUND;
}

// Stored copies of fingerprint (fq, fr): the reference walk finds the last
// match (rsqf_lookup_one); the copies are the equal remainders ending there,
// since remainders ascend inside a run.  A counting query over the multiset
// the filter keeps (the reference has no counters, README.md:74).
// ---------------------------------------------------------------- counting query
// Copies of a fingerprint, from run ranks rather than the 8-bit offsets: in a
// valid RSQF the k-th run end (from slot 0) closes the run of the k-th
// occupied quotient, so with P_occ / P_end, the per-block exclusive prefixes
// of the occupied and run-end bits (k_rc_pop + a row scan), the run of fq is
// found exactly however long the runs before it are:
//   k      = P_occ[b] + #(occupied bits of block b up to fq),
//   e      = slot of the k-th run end (binary search of P_end, then select),
//   start  = max(fq, slot of the (k-1)-th run end + 1),
// and the copies are the slots of [start, e] holding remainder fr.
__global__ __launch_bounds__(256) void k_rc_pop(const uint64_t* __restrict__ meta, uint32_t nb,
                                                uint32_t* __restrict__ cnt) {
  const uint32_t b = blockIdx.x * 256u + threadIdx.x;
  if (b >= nb) return;
  cnt[2 * (size_t)b] = (uint32_t)__popcll(meta[4 * (size_t)b]);
  cnt[2 * (size_t)b + 1] = (uint32_t)__popcll(meta[4 * (size_t)b + 1]);
}

// Slot of the k-th run end (k >= 1), searching blocks from b0 on; P: the
// exclusive prefixes (occupied, run end) per block; tot_end: all run ends.
__device__ __forceinline__ uint64_t rc_run_end(const RsqfDev& f, const uint32_t* __restrict__ P, uint32_t b0,
                                               uint32_t k) {
  uint32_t lo = b0, hi = f.nb;  // the block: the last with P_end < k
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (P[2 * (size_t)mid + 1] < k) lo = mid;
    else hi = mid;
  }
  return (uint64_t)lo * 64 + select_bit(f.ends(lo), k - P[2 * (size_t)lo + 1]);
}

template <int R>
__device__ __forceinline__ uint32_t rsqf_count_exact(const RsqfDev& f, const uint32_t* __restrict__ P,
                                                     uint32_t tot_end, uint32_t fq, uint32_t fr) {
  const uint32_t b = fq >> 6, i = fq & 63u;
  const uint64_t occ = f.occ(b);
  if (!((occ >> i) & 1ull)) return 0u;
  const uint32_t k = P[2 * (size_t)b] + (uint32_t)__popcll(occ & (i == 63 ? ~0ull : ((2ull << i) - 1ull)));
  if (k > tot_end) return 0u;  // not a valid table: no run end for this quotient
  const uint64_t e = rc_run_end(f, P, b, k);
  uint64_t start = fq;
  if (k > 1) {
    // the previous run end lies before e: search from block 0 up to e's block
    uint32_t lo = 0, hi = (uint32_t)(e >> 6) + 1;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (P[2 * (size_t)mid + 1] < k - 1) lo = mid;
      else hi = mid;
    }
    const uint64_t pe = (uint64_t)lo * 64 + select_bit(f.ends(lo), k - 1 - P[2 * (size_t)lo + 1]);
    if (pe + 1 > start) start = pe + 1;
  }
  uint32_t c = 0;
  for (uint64_t s = start; s <= e; ++s) c += f.get<R>((uint32_t)(s >> 6), (uint32_t)(s & 63u)) == fr;
  return c;
}

template <int R>
__global__ __launch_bounds__(256) void k_rsqf_count(RsqfDev f, uint32_t q, const uint32_t* __restrict__ P,
                                                    const uint32_t* __restrict__ totals,
                                                    const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint64_t fp = fingerprint(keys[i], q + f.r);
  out[i] = rsqf_count_exact<R>(f, P, totals[1], (uint32_t)((fp >> f.r) & mask64(q)),
                               (uint32_t)(fp & ((1u << f.r) - 1u)));
}

template <int R>
__global__ __launch_bounds__(256) void k_rsqf_lookup(RsqfDev f, uint32_t q, const uint32_t* __restrict__ keys,
                                                     uint32_t n, uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint64_t fp = fingerprint(keys[i], q + f.r);
  out[i] = rsqf_lookup_one<R>(f, (uint32_t)((fp >> f.r) & mask64(q)), (uint32_t)(fp & ((1u << f.r) - 1u)));
}

// ---------------------------------------------------------------- decode
// In a valid RSQF the entries in slot order are the stored fingerprints in
// ascending order, and the k-th run end (counting from slot 0) closes the run
// of the k-th occupied quotient.  Slot s therefore holds an entry iff the
// quotient Q of run k = 1 + #(run ends before s) satisfies Q <= s, and the
// entry is Q << r | remainder(s).  Slots are decoded in windows of RD_SLOTS
// (64 blocks):
//   k_rd_pop      occupied and run-end counts per window,
//   k_scan_rows   their exclusive prefixes O_w, E_w,
//   k_rd_count    entries per window,
//   k_scan_rows   output offsets,
//   k_rd_emit     the fingerprints, staged in LDS, written coalesced.
// Window w needs the quotients of runs E_w+1 ... E_w+1+#(run ends in w): a
// list built in LDS from the occupied words, starting at the block that holds
// the (E_w+1)-th occupied quotient (found through O).  Replaces the
// reference's sequential walk of runs (RSQF.cpp:314-569 reads them back the
// same way, one run at a time).
constexpr uint32_t RD_BLOCKS = 64;
constexpr uint32_t RD_SLOTS = RD_BLOCKS * 64;
constexpr uint32_t RD_THREADS = 256;  // 16 slots per thread
constexpr uint32_t RD_LIST = RD_SLOTS + 1;

// hb (may be NULL): the bounds of a packed decode, reset here (nhv of them
// counted by k_rd_emit, the others never reached).
__global__ __launch_bounds__(256) void k_rd_pop(const uint64_t* __restrict__ meta, uint32_t nb, uint32_t W,
                                                uint32_t* __restrict__ wp, uint32_t* __restrict__ hb, uint32_t nhv) {
  const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (hb && blockIdx.x == 0 && threadIdx.x < 3) hb[threadIdx.x] = threadIdx.x < nhv ? 0u : 0xFFFFFFFFu;
  if (w >= W) return;
  const uint32_t b = w * RD_BLOCKS + lane;
  uint32_t o = 0, e = 0;
  if (b < nb) {
    const uint4 x = reinterpret_cast<const uint4*>(meta)[2 * (size_t)b];
    o = (uint32_t)__popc(x.x) + (uint32_t)__popc(x.y);
    e = (uint32_t)__popc(x.z) + (uint32_t)__popc(x.w);
  }
  o = wave_incl_add(o);
  e = wave_incl_add(e);
  if (lane == 63) {
    wp[2 * (size_t)w] = o;
    wp[2 * (size_t)w + 1] = e;
  }
}

// s_list[i] = quotient of the (E_w + 1 + i)-th occupied slot for i < RD_LIST,
// RS_NOT_FOUND where that quotient lies past the window.
__device__ void rd_build_list(const RsqfDev& f, const uint32_t* __restrict__ wp, uint32_t w, uint32_t* s_list,
                              uint32_t* s_tmp4, uint32_t* s_v) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < RD_LIST; i += RD_THREADS) s_list[i] = RS_NOT_FOUND;
  const uint32_t R0 = wp[2 * (size_t)w + 1] + 1;
  // window v holding the R0-th occupied quotient: the last one with O_v < R0
  // (O_0 = 0, so there is one).  Usually w itself or a near predecessor: the
  // block checks the 256 windows before w at once, thread t window w-1-t,
  // and only a run longer than that falls back to a binary search.
  if (wp[2 * (size_t)w] < R0) {
    if (tid == 0) {
      s_v[0] = w;
      s_v[1] = wp[2 * (size_t)w];
    }
  } else {
    if (tid == 0) s_v[0] = 0xFFFFFFFFu;
    __syncthreads();
    if (tid < w && wp[2 * (size_t)(w - 1 - tid)] < R0) atomicMin(&s_v[0], tid);
    __syncthreads();
    if (tid == 0) {
      uint32_t v;
      if (s_v[0] != 0xFFFFFFFFu) {
        v = w - 1 - s_v[0];
      } else {
        uint32_t lo = 0, hi = w - RD_THREADS;  // O_lo < R0 <= O_hi
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (wp[2 * (size_t)mid] < R0) lo = mid;
          else hi = mid;
        }
        v = lo;
      }
      s_v[0] = v;
      s_v[1] = wp[2 * (size_t)v];
    }
  }
  __syncthreads();
  const uint32_t v = s_v[0];
  uint32_t base = s_v[1];  // occupied quotients before the current window
  // one window of 64 blocks per step: thread t holds 16 occupied bits of
  // block t/4, ranks them by a block scan and places each set bit's quotient
  for (uint32_t x = v; x <= w && base < R0 + RD_LIST - 1; ++x) {
    const uint32_t b = x * RD_BLOCKS + (tid >> 2), sh = 16 * (tid & 3);
    const uint32_t o16 = b < f.nb ? (uint32_t)(f.occ(b) >> sh) & 0xFFFFu : 0u;
    uint32_t tot;
    uint32_t g = base + block256_excl_scan((uint32_t)__popc(o16), s_tmp4, &tot) + 1;  // rank of the first set bit
    if (g + (uint32_t)__popc(o16) > R0 && g < R0 + RD_LIST) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        if ((o16 >> j) & 1u) {
          if (g >= R0 && g - R0 < RD_LIST) s_list[g - R0] = b * 64 + sh + (uint32_t)j;
          ++g;
        }
      }
    }
    base += tot;
  }
  __syncthreads();
}

// This thread's 16 slots: run ends (bit j = slot s0 + j) and the exclusive
// count of run ends before them in the window.
__device__ __forceinline__ uint32_t rd_ends16(const RsqfDev& f, uint32_t w, uint32_t* s0) {
  const uint32_t t = threadIdx.x;
  const uint32_t b = w * RD_BLOCKS + (t >> 2);
  *s0 = w * RD_SLOTS + 16 * t;
  return b < f.nb ? (uint32_t)(f.ends(b) >> (16 * (t & 3))) & 0xFFFFu : 0u;
}

// Bit j set: slot s0 + j holds an entry (its quotient in q[j]).
__device__ __forceinline__ uint32_t rd_slots(const RsqfDev& f, uint32_t e16, uint32_t rank0, uint32_t s0,
                                             const uint32_t* s_list, uint32_t* qv) {
  uint32_t m = 0;
  const bool in_table = (s0 >> 6) < f.nb;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint32_t tr = rank0 + (uint32_t)__popc(e16 & ((1u << j) - 1u));
    const uint32_t Q = tr < RD_LIST ? s_list[tr] : RS_NOT_FOUND;
    qv[j] = Q;
    if (in_table && Q <= s0 + (uint32_t)j) m |= 1u << j;
  }
  return m;
}

// Entries per window without the quotient list: slot s holds an entry iff
// the k-th occupied quotient (k = 1 + run ends before s) is <= s, i.e. iff
// D(s) = #(occupied bits <= s) - #(run ends < s) > 0 -- the runs opened up
// to s and not yet closed.  One block scan of the two popcounts per thread
// (packed in the halves of a word: a window has 4096 slots).
__global__ __launch_bounds__(RD_THREADS) void k_rd_count(RsqfDev f, const uint32_t* __restrict__ wp,
                                                         uint32_t* __restrict__ wc) {
  __shared__ uint32_t s_tmp4[4];
  const uint32_t w = blockIdx.x, t = threadIdx.x;
  const uint32_t b = w * RD_BLOCKS + (t >> 2), sh = 16 * (t & 3);
  uint32_t o16 = 0, e16 = 0;
  if (b < f.nb) {
    o16 = (uint32_t)(f.occ(b) >> sh) & 0xFFFFu;
    e16 = (uint32_t)(f.ends(b) >> sh) & 0xFFFFu;
  }
  const uint32_t pre = block256_excl_scan(((uint32_t)__popc(o16) << 16) | (uint32_t)__popc(e16), s_tmp4, nullptr);
  int32_t D = (int32_t)(wp[2 * (size_t)w] - wp[2 * (size_t)w + 1]) + (int32_t)(pre >> 16) - (int32_t)(pre & 0xFFFFu);
  uint32_t n = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    D += (int32_t)((o16 >> j) & 1u);
    n += D > 0 ? 1u : 0u;
    D -= (int32_t)((e16 >> j) & 1u);
  }
  if (b >= f.nb) n = 0;
  uint32_t total;
  (void)block256_excl_scan(n, s_tmp4, &total);
  if (t == 0) wc[w] = total;
}

// PK: out receives the low 32 bits of each 64-bit fingerprint (so the LDS
// stage holds 32-bit words too), and the window holding quotient
// Qk = k << (32 - r) (k = 1 .. nhv) -- or a later one whose first entries
// still have smaller quotients -- adds its entries of quotient below Qk to
// hb[k - 1], which ends as the index where the high word reaches k.
template <typename K, int R, bool PK>
__global__ __launch_bounds__(RD_THREADS) void k_rd_emit(RsqfDev f, const uint32_t* __restrict__ wp,
                                                        const uint32_t* __restrict__ wc, K* __restrict__ out,
                                                        uint32_t* __restrict__ hb, uint32_t nhv) {
  // the quotient list, then (after every thread has read it) the window's entries
  using EK = typename std::conditional<PK, uint32_t, K>::type;
  __shared__ __align__(16) EK s_buf[RD_SLOTS * sizeof(EK) >= RD_LIST * 4 ? RD_SLOTS : RD_LIST];
  __shared__ uint32_t s_tmp4[4];
  __shared__ uint32_t s_v[2];
  __shared__ uint32_t s_fq;  // PK: quotient of the window's first entry
  uint32_t* s_list = reinterpret_cast<uint32_t*>(s_buf);
  const uint32_t w = blockIdx.x;
  const uint32_t rr = R ? (uint32_t)R : f.r;
  // this thread's run ends and (8- and 16-bit remainders) remainder words are
  // loaded first, so that their latency overlaps the list build
  uint32_t s0;
  const uint32_t e16 = rd_ends16(f, w, &s0);
  const uint32_t b = s0 >> 6;
  uint4 x0 = make_uint4(0u, 0u, 0u, 0u), x1 = x0;
  if (b < f.nb) {
    if (R == 8) {
      x0 = *reinterpret_cast<const uint4*>(f.rem + s0);
    } else if (R == 16) {
      const uint4* p = reinterpret_cast<const uint4*>(f.rem + 2 * (size_t)s0);
      x0 = p[0];
      x1 = p[1];
    }
  }
  rd_build_list(f, wp, w, s_list, s_tmp4, s_v);
  uint32_t tot;
  const uint32_t rank0 = block256_excl_scan((uint32_t)__popc(e16), s_tmp4, &tot);
  uint32_t qv[16];
  const uint32_t m = rd_slots(f, e16, rank0, s0, s_list, qv);
  // remainders of the 16 slots
  uint32_t rem[16];
  if (R == 8) {
    const uint32_t wd[4] = {x0.x, x0.y, x0.z, x0.w};
#pragma unroll
    for (int j = 0; j < 16; ++j) rem[j] = (wd[j >> 2] >> (8 * (j & 3))) & 0xFFu;
  } else if (R == 16) {
    const uint32_t wd[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int j = 0; j < 16; ++j) rem[j] = (wd[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
  } else if (m) {
#pragma unroll
    for (int j = 0; j < 16; ++j) rem[j] = (m >> j) & 1u ? f.get<0>(b, (s0 & 63) + (uint32_t)j) : 0u;
  }
  uint32_t cnt;
  const uint32_t pos = block256_excl_scan((uint32_t)__popc(m), s_tmp4, &cnt);
  __syncthreads();  // every thread is done with s_list
  uint32_t p = pos;
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if ((m >> j) & 1u) s_buf[p++] = (EK)(((K)qv[j] << rr) | (K)rem[j]);
  if (PK && m && pos == 0) s_fq = qv[__builtin_ctz(m)];  // the thread holding the window's first entry
  __syncthreads();
  if constexpr (PK) {
    uint32_t* dst = reinterpret_cast<uint32_t*>(out) + wc[w];
    for (uint32_t i = threadIdx.x; i < cnt; i += RD_THREADS) dst[i] = s_buf[i];
    // (w, cnt and s_fq are the same in every thread: the branches are uniform)
    for (uint32_t k = 1; k <= nhv; ++k) {
      const uint64_t Qk = (uint64_t)k << (32 - rr);
      const uint64_t wk = Qk / RD_SLOTS;  // window of quotient Qk
      if (w < wk) break;
      if (w > wk && (cnt == 0 || s_fq >= Qk)) continue;
      uint32_t c = 0;  // this thread's entries of quotient below Qk
#pragma unroll
      for (int j = 0; j < 16; ++j) c += ((m >> j) & 1u) && qv[j] < Qk ? 1u : 0u;
      uint32_t below;
      (void)block256_excl_scan(c, s_tmp4, &below);
      const uint32_t add = below + (w == wk ? wc[w] : 0u);
      if (threadIdx.x == 0 && add) atomicAdd(&hb[k - 1], add);
    }
  } else {
    K* dst = out + wc[w];
    for (uint32_t i = threadIdx.x; i < cnt; i += RD_THREADS) dst[i] = s_buf[i];
  }
}

// ---------------------------------------------------------------- rebuild
// Sorted fingerprints (build; the merge of the decoded table and a sorted
// batch for insert; the decoded table with keep flags for delete) to the
// canonical layout, through the chunked location scan (k_rr_agg, rb_scan)
// and a writer that walks each sub-tile again on top of its prefix.
//
// The writer runs on a table cleared beforehand and leaves every field with
// one owner:
//   remainders   of slots [L_lo, L_hi) -- the sub-tile's own locations;
//   run ends     of slots [X.l, Y.l) -- each run head marks the end of the
//                previous kept entry's run (the last run end: k_rr_tail);
//   occupieds    of its quotient blocks -- the first and last may be shared
//                with the neighbouring sub-tiles;
//   offset_j     by the first kept entry with quotient >= 64 j (the crossing
//                entry; blocks after the last quotient: k_rr_tail).
// A sub-tile assembles its fields in an LDS window of RW_WB blocks and
// writes each field of a block with plain stores where it owns it whole,
// atomic ORs where it shares it.  Replaces the write stage of insertGPU
// (RSQF.cu:882-988) with the end state of inserting the fingerprints in
// ascending order (RSQF.cpp:314-400).
constexpr uint32_t RW_WB = 128;  // blocks per LDS window (8192 slots)
constexpr uint32_t RR_MERGE_PER = 2;  // sub-tiles per workgroup in merge rebuilds (the next one staged)

// Merge rebuilds stage a sub-tile's inputs in LDS with global->LDS loads,
// which take no VGPRs and run on while the workgroup processes the previous
// sub-tile: 32-bit words, the low words in s_lo and, for 64-bit values, the
// high words in s_hi (a packed A takes its high words from the index).
// Positions [0, a1 - a0) hold A[a0 ..], the rest B[d0 - a0 ..].
template <typename K, bool PK>
__device__ __forceinline__ void rr_stage_issue(const FpArray<K, PK>& A, const K* __restrict__ B, uint32_t a0,
                                               uint32_t a1, uint32_t d0, uint32_t tn, uint32_t* s_lo,
                                               uint32_t* s_hi) {
  constexpr bool WIDE = sizeof(K) == 8;
  const uint32_t la = a1 - a0, b0 = d0 - a0;
  const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t* A32 = reinterpret_cast<const uint32_t*>(A.p);
  const uint32_t* B32 = reinterpret_cast<const uint32_t*>(B);
#pragma unroll
  for (int k = 0; k < RB_ITEMS; ++k) {
    // wave wid, step k: the 64 consecutive positions from base (lane l -> base + l)
    const uint32_t base = ((uint32_t)k * RB_WAVES + wid) * 64;
    const uint32_t pos = base + lane;
    if (pos < tn) {
      const bool inA = pos < la;
      const uint64_t w = inA ? (uint64_t)a0 + pos : (uint64_t)b0 + (pos - la);  // entry index
      const bool two = WIDE && (!PK || !inA);                                   // 64-bit words in memory
      const uint32_t* src = (inA ? A32 : B32) + (two ? 2 * w : w);
      __builtin_amdgcn_global_load_lds(src, s_lo + base, 4, 0, 0);
      if (two) __builtin_amdgcn_global_load_lds(src + 1, s_hi + base, 4, 0, 0);
    }
  }
}

// Views of a staged sub-tile's A and B parts.
template <typename K, bool PK>
struct StagedA {
  const uint32_t* lo;
  const uint32_t* hi;
  FpArray<K, PK> A;
  uint32_t a0;
  __device__ __forceinline__ K operator[](uint32_t i) const {
    if constexpr (PK) return A.widen(a0 + i, lo[i]);
    else if constexpr (sizeof(K) == 8) return (K)(((uint64_t)hi[i] << 32) | lo[i]);
    else return (K)lo[i];
  }
};
template <typename K>
struct StagedB {
  const uint32_t* lo;
  const uint32_t* hi;
  __device__ __forceinline__ K operator[](uint32_t j) const {
    if constexpr (sizeof(K) == 8) return (K)(((uint64_t)hi[j] << 32) | lo[j]);
    else return (K)lo[j];
  }
};

// This thread's RB_ITEMS entries of the stable merge (A first on ties) of a
// staged sub-tile with A range [a0, a1) and tn entries.  A batch adds few
// entries to a sub-tile, so instead of a merge-path search per thread every
// B entry finds its output position (its index + the A entries <= it) and
// marks it in the bitmap s_bmp (RB_TILE bits, zero on entry); a prefix count
// of the bitmap words (s_wpre) then gives each thread its B and A offsets.
// Every thread of the block calls it (two barriers inside).
template <typename K, bool PK>
__device__ __forceinline__ void rr_merge_staged(const FpArray<K, PK>& A, const uint32_t* s_lo, const uint32_t* s_hi,
                                                uint32_t a0, uint32_t a1, uint32_t tn, K* v, uint32_t* s_bmp,
                                                uint32_t* s_wpre) {
  const uint32_t la = a1 - a0, lb = tn - la;
  const StagedA<K, PK> sA{s_lo, s_hi, A, a0};
  const StagedB<K> sB{s_lo + la, s_hi + la};
  for (uint32_t j = threadIdx.x; j < lb; j += RB_THREADS) {
    const K x = sB[j];
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
  const uint32_t bits = w >> (i0 & 31);
#pragma unroll
  for (int j = 0; j < RB_ITEMS; ++j) {
    K x = 0;
    if (i0 + j < tn) x = ((bits >> j) & 1u) ? sB[bi++] : sA[ai++];
    v[j] = x;
  }
}

// This thread's RB_ITEMS consecutive entries of sub-tile [d0, d0+tn) of A
// (no merge) and, under KEEP_FLAGS, their flags.
template <typename K, int MODE, bool PK>
__device__ __forceinline__ void rr_load(const FpArray<K, PK>& A, const uint8_t* __restrict__ flags, uint32_t d0,
                                        uint32_t tn, K* v, uint32_t& fl) {
  const int tid = threadIdx.x;
  const uint32_t i0 = (uint32_t)tid * RB_ITEMS;
  fl = 0;
  {
    const uint64_t g0 = (uint64_t)d0 + i0;
    if (i0 + RB_ITEMS <= tn) {
      // d0 is a multiple of RB_TILE and A is 256-byte aligned
      constexpr int PER = PK ? 4 : 16 / (int)sizeof(K);
      const uint4* p = reinterpret_cast<const uint4*>(reinterpret_cast<const K*>(A.p) + g0);
      if (PK) p = reinterpret_cast<const uint4*>(reinterpret_cast<const uint32_t*>(A.p) + g0);
#pragma unroll
      for (int k = 0; k < RB_ITEMS / PER; ++k) {
        const uint4 x = p[k];
        if (PK) {
          const uint32_t i = (uint32_t)g0 + 4 * k;
          v[4 * k] = A.widen(i, x.x);
          v[4 * k + 1] = A.widen(i + 1, x.y);
          v[4 * k + 2] = A.widen(i + 2, x.z);
          v[4 * k + 3] = A.widen(i + 3, x.w);
        } else if (sizeof(K) == 8) {
          v[2 * k] = (K)(((uint64_t)x.y << 32) | x.x);
          v[2 * k + 1] = (K)(((uint64_t)x.w << 32) | x.z);
        } else {
          v[4 * k] = (K)x.x;
          v[4 * k + 1] = (K)x.y;
          v[4 * k + 2] = (K)x.z;
          v[4 * k + 3] = (K)x.w;
        }
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
        v[k] = (i0 + k < tn) ? A[(uint32_t)g0 + k] : (K)0;
        if (MODE == KEEP_FLAGS && i0 + k < tn && flags[g0 + k]) fl |= 1u << k;
      }
    }
  }
}

// FULL: tn == RB_TILE (no per-entry range tests).
template <typename K, int MODE, bool FULL = false>
__device__ __forceinline__ uint32_t rr_classify(const K* v, uint32_t fl, uint32_t tn, uint32_t r, uint32_t mq,
                                                RS& agg) {
  uint32_t keepbits = 0;
  agg = rs_none();
  const uint32_t i0 = (uint32_t)threadIdx.x * RB_ITEMS;
#pragma unroll
  for (int j = 0; j < RB_ITEMS; ++j) {
    if ((FULL || i0 + j < tn) && (MODE == KEEP_ALL || ((fl >> j) & 1u))) {
      keepbits |= 1u << j;
      rs_push(agg, (uint32_t)(v[j] >> r) & mq);
    }
  }
  return keepbits;
}

// PK: A is packed (FpArray) with its bounds at hb.
template <typename K, int MODE, bool MERGE, bool PK>
__global__ __launch_bounds__(RB_THREADS, 8) void k_rr_agg(const K* __restrict__ Araw, uint32_t na,
                                                          const uint32_t* __restrict__ na_dev,
                                                          const K* __restrict__ B, uint32_t nbk,
                                                          const uint8_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ splits, uint32_t q, uint32_t r,
                                                          uint32_t chunk, uint4* __restrict__ agg,
                                                          const uint32_t* __restrict__ hb, uint32_t* __restrict__ ctr) {
  // the rebuild's counters start at zero (nothing reads them before the scan)
  if (blockIdx.x == 0 && threadIdx.x >= CTR_OP && threadIdx.x < CTR_TOTAL) ctr[threadIdx.x] = 0u;
  __shared__ uint32_t s_lo[MERGE ? RB_TILE : 1];
  __shared__ uint32_t s_hi[MERGE && sizeof(K) == 8 ? RB_TILE : 1];
  __shared__ uint32_t s_bmp[MERGE ? RB_TILE / 32 : 1], s_wpre[MERGE ? RB_TILE / 32 : 1];
  __shared__ uint4 s_wave[RB_WAVES + 1];
  const FpArray<K, PK> A = fp_array<K, PK>(Araw, hb);
  if (MERGE && threadIdx.x < RB_TILE / 32) s_bmp[threadIdx.x] = 0u;
  if (na_dev) na = *na_dev;  // stored count on the device (the grid covers an upper bound)
  const uint64_t N = (uint64_t)na + (MERGE ? nbk : 0u);
  const uint32_t mq = (uint32_t)mask64(q);
  const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t c1 = c0 + chunk < N ? c0 + chunk : N;
  RS X = rs_none();
  // MERGE: splits of the current sub-tile (ca0, ca1) and the next one's end
  // (na1), loaded a sub-tile ahead
  uint32_t ca0 = 0, ca1 = 0, na1 = 0;
  if (MERGE && c0 < c1) {
    const uint32_t u0 = (uint32_t)(c0 / RB_TILE);
    ca0 = splits[u0];
    ca1 = splits[u0 + 1];
    if (c0 + RB_TILE < c1) na1 = splits[u0 + 2];
    rr_stage_issue<K, PK>(A, B, ca0, ca1, (uint32_t)c0, (uint32_t)(c1 - c0 < RB_TILE ? c1 - c0 : RB_TILE), s_lo,
                          s_hi);
  }
  for (uint64_t d0 = c0; d0 < c1; d0 += RB_TILE) {
    const uint32_t tn = (uint32_t)(c1 - d0 < (uint64_t)RB_TILE ? c1 - d0 : (uint64_t)RB_TILE);
    K v[RB_ITEMS];
    uint32_t fl = 0;
    if (MERGE) {
      const uint32_t u = (uint32_t)(d0 / RB_TILE);
      const uint64_t dn = d0 + RB_TILE;
      rb_stage_wait();
      __syncthreads();
      rr_merge_staged<K, PK>(A, s_lo, s_hi, ca0, ca1, tn, v, s_bmp, s_wpre);
      const uint32_t nn1 = dn + RB_TILE < c1 ? splits[u + 3] : 0u;
      __syncthreads();  // every thread holds its entries: the stage and the bitmap are free
      if (threadIdx.x < RB_TILE / 32) s_bmp[threadIdx.x] = 0u;
      if (dn < c1)
        rr_stage_issue<K, PK>(A, B, ca1, na1, (uint32_t)dn, (uint32_t)(c1 - dn < RB_TILE ? c1 - dn : RB_TILE), s_lo,
                              s_hi);
      ca0 = ca1;
      ca1 = na1;
      na1 = nn1;
    } else {
      rr_load<K, MODE, PK>(A, flags, (uint32_t)d0, tn, v, fl);
    }
    RS a;
    if (tn == RB_TILE) (void)rr_classify<K, MODE, true>(v, fl, tn, r, mq, a);
    else (void)rr_classify<K, MODE, false>(v, fl, tn, r, mq, a);
    RS ex;
    X = rs_combine(X, rs_block(a, &ex, s_wave));
    __syncthreads();
  }
  if (threadIdx.x == 0) agg[blockIdx.x] = rs_pack(X);
}

// Zeroes the table (meta and remainders), unless the scan raised a capacity error.
__global__ __launch_bounds__(256) void k_rr_clear(uint8_t* __restrict__ t, uint64_t bytes,
                                                  const uint32_t* __restrict__ ctr) {
  if (ctr[CTR_ERR]) return;
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  const uint64_t n16 = bytes / 16;
  uint4* t16 = reinterpret_cast<uint4*>(t);
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256) t16[i] = z;
  if (blockIdx.x == 0 && threadIdx.x < bytes - 16 * n16) t[16 * n16 + threadIdx.x] = 0;
}

// One kept entry's contributions to the window [jw0, jw1) (block indices).
// prev / lastf: location and quotient of the previous kept entry (-1 /
// 0xFFFFFFFF when there is none).
template <int R>
__device__ __forceinline__ void rr_visit(uint32_t fq, uint32_t fr, int32_t lc, int32_t prev, uint32_t lastf,
                                         uint32_t r, uint32_t jw0, uint32_t jw1, uint32_t* s_occ, uint32_t* s_end,
                                         uint32_t* s_off, uint32_t* s_rem) {
  const uint32_t sw0 = jw0 * 64;
  const uint32_t lb = (uint32_t)lc >> 6;
  if (lb >= jw0 && lb < jw1) {
    const uint32_t s = (uint32_t)lc - sw0;
    if (R == 8) {
      reinterpret_cast<uint8_t*>(s_rem)[s] = (uint8_t)fr;
    } else if (R == 16) {
      reinterpret_cast<uint16_t*>(s_rem)[s] = (uint16_t)fr;
    } else {
      const uint32_t bit = s * r;
      atomicOr(&s_rem[bit >> 5], fr << (bit & 31));
      if ((bit & 31) + r > 32) atomicOr(&s_rem[(bit >> 5) + 1], fr >> (32 - (bit & 31)));
    }
  }
  if (fq != lastf) {
    const uint32_t qb = fq >> 6;
    if (qb >= jw0 && qb < jw1) atomicOr(&s_occ[2 * (qb - jw0) + ((fq >> 5) & 1u)], 1u << (fq & 31));
    if (prev >= 0) {
      const uint32_t pb = (uint32_t)prev >> 6;
      if (pb >= jw0 && pb < jw1)
        atomicOr(&s_end[2 * (pb - jw0) + (((uint32_t)prev >> 5) & 1u)], 1u << ((uint32_t)prev & 31));
      // blocks whose first slot lies in (lastf, min(fq, prev)]: offset = slots
      // of block j taken by entries of quotients < 64 j
      const uint32_t lim = fq < (uint32_t)prev ? fq : (uint32_t)prev;
      uint32_t j = (lastf >> 6) + 1;
      if (j < jw0) j = jw0;
      const uint32_t je = (lim >> 6) + 1 < jw1 ? (lim >> 6) + 1 : jw1;
      for (; j < je; ++j) s_off[j - jw0] = ((uint32_t)prev - 64 * j + 1) & 0xFFu;
    }
  }
}

template <typename K, int MODE, bool MERGE, int R, bool PK>
__global__ __launch_bounds__(RB_THREADS, 6) void k_rr_write(uint8_t* __restrict__ base, uint32_t nb, uint32_t q,
                                                            uint32_t r, const K* __restrict__ Araw, uint32_t na,
                                                            const uint32_t* __restrict__ na_dev,
                                                            const K* __restrict__ B, uint32_t nbk,
                                                            const uint8_t* __restrict__ flags,
                                                            const uint32_t* __restrict__ splits, uint32_t chunk,
                                                            const uint4* __restrict__ pre,
                                                            const uint32_t* __restrict__ ctr,
                                                            const uint32_t* __restrict__ hb) {
  if (ctr[CTR_ERR]) return;  // capacity error: the table stays as it was
  const FpArray<K, PK> A = fp_array<K, PK>(Araw, hb);
  __shared__ uint32_t s_lo[MERGE ? RB_TILE : 1];
  __shared__ uint32_t s_hi[MERGE && sizeof(K) == 8 ? RB_TILE : 1];
  __shared__ uint32_t s_bmp[MERGE ? RB_TILE / 32 : 1], s_wpre[MERGE ? RB_TILE / 32 : 1];
  if (MERGE && threadIdx.x < RB_TILE / 32) s_bmp[threadIdx.x] = 0u;
  __shared__ __align__(16) uint32_t s_rem[RW_WB * 32];  // 2r words per block, r <= 16
  __shared__ uint32_t s_occ[RW_WB * 2];
  __shared__ uint32_t s_end[RW_WB * 2];
  __shared__ uint32_t s_off[RW_WB];
  __shared__ uint4 s_wave[RB_WAVES + 1];
  __shared__ uint32_t s_b[2];
  // the window fields start cleared (the barriers before the first visit order it)
  for (uint32_t i = threadIdx.x; i < RW_WB * 32; i += RB_THREADS) s_rem[i] = 0u;
  for (uint32_t i = threadIdx.x; i < RW_WB; i += RB_THREADS) {
    s_occ[2 * i] = s_occ[2 * i + 1] = 0u;
    s_end[2 * i] = s_end[2 * i + 1] = 0u;
    s_off[i] = 0u;
  }
  const int tid = threadIdx.x;
  const uint32_t rr = R ? (uint32_t)R : r;
  const uint32_t mq = (uint32_t)mask64(q);
  const uint32_t rmask = (1u << rr) - 1u;
  if (na_dev) na = *na_dev;
  const uint64_t N = (uint64_t)na + (MERGE ? nbk : 0u);
  const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t c1 = c0 + chunk < N ? c0 + chunk : N;
  uint64_t* meta = reinterpret_cast<uint64_t*>(base);
  uint32_t* remg = reinterpret_cast<uint32_t*>(base + rsqf_meta_bytes(nb));
  const uint32_t rw = 2 * rr;  // remainder words per block
  RS X = rs_unpack(pre[blockIdx.x]);
  uint32_t ca0 = 0, ca1 = 0, na1 = 0;  // as in k_rr_agg
  if (MERGE && c0 < c1) {
    const uint32_t u0 = (uint32_t)(c0 / RB_TILE);
    ca0 = splits[u0];
    ca1 = splits[u0 + 1];
    if (c0 + RB_TILE < c1) na1 = splits[u0 + 2];
    rr_stage_issue<K, PK>(A, B, ca0, ca1, (uint32_t)c0, (uint32_t)(c1 - c0 < RB_TILE ? c1 - c0 : RB_TILE), s_lo,
                          s_hi);
  }
  for (uint64_t d0 = c0; d0 < c1; d0 += RB_TILE) {
    const uint32_t u = (uint32_t)(d0 / RB_TILE);
    const uint32_t tn = (uint32_t)(c1 - d0 < (uint64_t)RB_TILE ? c1 - d0 : (uint64_t)RB_TILE);
    K v[RB_ITEMS];
    uint32_t fl = 0;
    if (MERGE) {
      const uint64_t dn = d0 + RB_TILE;
      rb_stage_wait();
      __syncthreads();
      rr_merge_staged<K, PK>(A, s_lo, s_hi, ca0, ca1, tn, v, s_bmp, s_wpre);
      const uint32_t nn1 = dn + RB_TILE < c1 ? splits[u + 3] : 0u;
      __syncthreads();  // every thread holds its entries: the stage and the bitmap are free
      if (threadIdx.x < RB_TILE / 32) s_bmp[threadIdx.x] = 0u;
      if (dn < c1)
        rr_stage_issue<K, PK>(A, B, ca1, na1, (uint32_t)dn, (uint32_t)(c1 - dn < RB_TILE ? c1 - dn : RB_TILE), s_lo,
                              s_hi);
      ca0 = ca1;
      ca1 = na1;
      na1 = nn1;
    } else {
      rr_load<K, MODE, PK>(A, flags, (uint32_t)d0, tn, v, fl);
    }
    RS a;
    const uint32_t keepbits = tn == RB_TILE ? rr_classify<K, MODE, true>(v, fl, tn, rr, mq, a)
                                            : rr_classify<K, MODE, false>(v, fl, tn, rr, mq, a);
    RS ex;
    const RS sub = rs_block(a, &ex, s_wave);
    const RS me = rs_combine(X, ex);
    const RS Y = rs_combine(X, sub);
    if (sub.c > 0) {
      const int32_t prev0 = (int32_t)me.lp - 1;  // -1: no entry before
      const uint32_t lastf0 = me.fq1 - 1u;       // 0xFFFFFFFF: none
      // blocks this thread's contributions touch: [t_lo, t_hi]
      uint32_t t_lo = 0xFFFFFFFFu, t_hi = 0;
      if (keepbits) {
        const int j0 = __builtin_ctz(keepbits);
        const uint32_t fq = (uint32_t)(v[j0] >> rr) & mq;
        t_lo = fq >> 6;
        if (fq != lastf0 && prev0 >= 0) {
          const uint32_t pb = (uint32_t)prev0 >> 6;
          if (pb < t_lo) t_lo = pb;
          const uint32_t lim = fq < (uint32_t)prev0 ? fq : (uint32_t)prev0;
          const uint32_t oj = (lastf0 >> 6) + 1;
          if (oj <= (lim >> 6) && oj < t_lo) t_lo = oj;
        }
        int32_t prev = prev0;
#pragma unroll
        for (int j = 0; j < RB_ITEMS; ++j) {
          if (keepbits & (1u << j)) {
            const int32_t f = (int32_t)((uint32_t)(v[j] >> rr) & mq);
            prev = f > prev + 1 ? f : prev + 1;
          }
        }
        t_hi = (uint32_t)prev >> 6;
        if (ex.c == 0) {
          // the thread holding the sub-tile's first kept entry: its quotient,
          // prev0 and lastf0 are the smallest of the sub-tile, so its t_lo is
          // the sub-tile's first block
          s_b[0] = t_lo;
          s_b[1] = fq >> 6;
        }
      }
      __syncthreads();
      const uint32_t B0 = s_b[0];
      const uint32_t B1 = ((Y.lp - 1u) >> 6) + 1;
      const uint32_t qfirst = s_b[1];            // quotient blocks of the first and last
      const uint32_t qlast = (Y.fq1 - 1u) >> 6;  // kept entries: occupied words shared
      const uint32_t Llo = X.lp, Lhi = Y.lp;     // remainders: [Llo, Lhi)
      const int64_t Elo = (int64_t)X.lp - 1;     // run ends: [Elo, Ehi)
      const int64_t Ehi = (int64_t)Y.lp - 1;
      for (uint32_t jw0 = B0; jw0 < B1; jw0 += RW_WB) {
        const uint32_t jw1 = jw0 + RW_WB < B1 ? jw0 + RW_WB : B1;
        const uint32_t nblk = jw1 - jw0;
        // several windows (sparse table): skip those without contributions
        if (B1 - B0 > RW_WB && !__syncthreads_or(t_lo < jw1 && t_hi >= jw0)) continue;
        // (the window fields are all zeros here: cleared at the start and by
        // each copy-out below)
        {
          int32_t prev = prev0;
          uint32_t lastf = lastf0;
#pragma unroll
          for (int j = 0; j < RB_ITEMS; ++j) {
            if (keepbits & (1u << j)) {
              const uint32_t fq = (uint32_t)(v[j] >> rr) & mq;
              const int32_t lc = (int32_t)fq > prev + 1 ? (int32_t)fq : prev + 1;
              rr_visit<R>(fq, (uint32_t)v[j] & rmask, lc, prev, lastf, rr, jw0, jw1, s_occ, s_end, s_off, s_rem);
              prev = lc;
              lastf = fq;
            }
          }
        }
        __syncthreads();
        // metadata: one thread per block and field
        for (uint32_t i = tid; i < nblk; i += RB_THREADS) {
          const uint32_t jb = jw0 + i;
          const uint64_t occ = ((uint64_t)s_occ[2 * i + 1] << 32) | s_occ[2 * i];
          const uint64_t end = ((uint64_t)s_end[2 * i + 1] << 32) | s_end[2 * i];
          const uint32_t off = s_off[i];
          s_occ[2 * i] = s_occ[2 * i + 1] = 0u;
          s_end[2 * i] = s_end[2 * i + 1] = 0u;
          s_off[i] = 0u;
          uint64_t* mrec = meta + 4 * (size_t)jb;
          if (occ) {
            if (jb == qfirst || jb == qlast) atomicOr((unsigned long long*)&mrec[0], (unsigned long long)occ);
            else mrec[0] = occ;
          }
          if (end) {
            const bool own = 64 * (int64_t)jb >= Elo && 64 * (int64_t)jb + 64 <= Ehi;
            if (own) mrec[1] = end;
            else atomicOr((unsigned long long*)&mrec[1], (unsigned long long)end);
          }
          if (off) mrec[2] = off;
        }
        // remainders: whole blocks owned by this sub-tile are stored, shared
        // ones ORed word by word
        for (uint32_t i = tid; i < nblk * rw; i += RB_THREADS) {
          const uint32_t jb = jw0 + i / rw;
          const uint32_t x = s_rem[i];
          s_rem[i] = 0u;
          const bool own = 64u * jb >= Llo && 64u * jb + 64 <= Lhi;
          uint32_t* g = remg + (size_t)jw0 * rw + i;
          if (own) *g = x;
          else if (x) atomicOr(g, x);
        }
        __syncthreads();
      }
    }
    X = Y;
    __syncthreads();
  }
}

// The last run end and the offsets of the blocks after the last quotient,
// from the scan total (last location lp - 1, last quotient fq1 - 1).
__global__ __launch_bounds__(64) void k_rr_tail(uint8_t* __restrict__ base, const uint4* __restrict__ total,
                                                const uint32_t* __restrict__ ctr) {
  if (ctr[CTR_ERR]) return;
  const RS t = rs_unpack(*total);
  if (t.c == 0) return;
  uint64_t* meta = reinterpret_cast<uint64_t*>(base);
  const uint32_t L = t.lp - 1u, F = t.fq1 - 1u;
  if (threadIdx.x == 0) atomicOr((unsigned long long*)&meta[4 * (size_t)(L >> 6) + 1], 1ull << (L & 63));
  for (uint32_t j = (F >> 6) + 1 + threadIdx.x; j <= (L >> 6); j += 64)
    meta[4 * (size_t)j + 2] = (L - 64 * j + 1) & 0xFFu;
}

// ---------------------------------------------------------------- windowed insert
// An insert that rewrites the table window by window in place instead of
// decoding it to a sorted array and rebuilding.  Window w is RD_BLOCKS
// blocks: the quotients and the slots [4096 w, 4096 w + 4096).
//   E_w  the stored entries of the window's quotients (the runs of its
//        occupied quotients, which may end in window w + 1);
//   B_w  the batch's fingerprints with those quotients (B is sorted);
//   M_w  their stable merge, stored entries first on ties (as the rebuild's
//        merge of the decoded table with the batch).
// The canonical layout of the whole is the location scan over the windows'
// merged entries in window order, so:
//   k_rd_pop + scan  occupied and run-end counts before each window: d_w =
//                    the runs of earlier quotients still open at its first
//                    slot (their ends are its first d_w run ends);
//   k_wm_bounds      each window's range of B;
//   k_wm_sum         each window's scan state, from its entries' quotient
//                    counts (run ends and occupieds only) and B_w's;
//   rb_scan          the state before each window: its incoming location;
//   k_wm_check       the blocks each window hands over (below);
//   k_wm_save        ... copied aside and cleared;
//   k_wm_write       each window decodes E_w, merges B_w, lays M_w out on top
//                    of its incoming state and writes its blocks.
// Decoding follows the rule of k_rd_emit: image slot s lies in run
// k = 1 + (run ends before s); the window's j-th run (k = d_w + j) is that
// of its j-th occupied quotient Q, and s holds an entry iff Q <= s.
//   In place, the first blocks of window x can be read and written by two
// workgroups: window x - 1's new entries land there (new incoming location)
// and its old runs ended there (old incoming location).  Below
// h_x = max(old, new incoming location of x) -- at most WM_SIDE blocks --
// k_wm_save copies the blocks' run ends and remainders aside and clears
// them; k_wm_write reads those blocks from the copy and ORs into them, while
// every other block has one reader and one writer, its own window, which
// stores it whole.  Occupieds and offsets belong to their window alone.
//   A window with more than WM_CAP merged entries, a run end that is not in
// its two-window image, or more than WM_SIDE blocks to hand over sends the
// call to the decode-and-rebuild path: the flag ctr[CTR_MISC] is raised
// before anything is written, and every writer checks it.
constexpr uint32_t WM_SIDE = 8;                     // blocks a window hands over at most (512 slots)
constexpr uint32_t WM_WAVES = 8;                    // k_wm_write: 16 image slots per thread
constexpr uint32_t WM_THREADS = 64 * WM_WAVES;
constexpr uint32_t WM_ITEMS = 16;                   // merged entries per thread
constexpr uint32_t WM_CAP = WM_THREADS * WM_ITEMS;  // 8192 merged entries per window
// image: windows w and w + 1 (128 blocks, 16 slots per thread); the
// window's entries end before block RD_BLOCKS + WM_SIDE (k_wm_check)
constexpr uint32_t WM_IMG = WM_THREADS / 4;
static_assert(WM_IMG >= RD_BLOCKS + WM_SIDE, "the image holds every slot a window's entries reach");
constexpr uint32_t WS_THREADS = 256;                // k_wm_sum: 16 quotients, 32 image slots per thread

// n copies of quotient fq appended to a scan state (rs_push n times).
__device__ __forceinline__ void rs_push_run(RS& s, uint32_t fq, uint32_t n) {
  s.lp = add_sat(s.lp > fq ? s.lp : fq, n);
  s.c += n;
  s.fq1 = fq + 1u;
}

// Entries per window-relative quotient, 16 bits each (a window that is not
// flagged holds fewer than 2^16 entries of one quotient).
__device__ __forceinline__ void wm_cnt_add(uint32_t* s_cnt, uint32_t p, uint32_t n) {
  atomicAdd(&s_cnt[p >> 1], n << (16 * (p & 1)));
}
__device__ __forceinline__ uint32_t wm_cnt_get(const uint32_t* s_cnt, uint32_t p) {
  return (s_cnt[p >> 1] >> (16 * (p & 1))) & 0xFFFFu;
}

// Each window's range of B: bstart[x] = first fingerprint with quotient >=
// RD_SLOTS x (x = 0 .. W), 8 lanes per window (group8_search).  Also resets
// the counters the insert uses.
constexpr uint32_t WB_LANES = 8;
template <typename K>
__global__ __launch_bounds__(256) void k_wm_bounds(const K* __restrict__ B, uint32_t n, uint32_t r, uint32_t W,
                                                   uint32_t* __restrict__ bstart, uint32_t* __restrict__ ctr) {
  if (blockIdx.x == 0 && threadIdx.x >= CTR_OP && threadIdx.x < CTR_TOTAL) ctr[threadIdx.x] = 0u;
  const uint32_t gt = blockIdx.x * 256 + threadIdx.x;
  const uint32_t x = gt / WB_LANES;
  // every lane stays for the ballots; groups past W search an empty range
  const uint64_t key = ((uint64_t)x * RD_SLOTS) << r;
  const uint32_t b = group8_search(0u, x <= W ? n : 0u, [&](uint32_t m) { return (uint64_t)B[m] < key; });
  if (gt % WB_LANES == 0 && x <= W) bstart[x] = b;
}

// Window w's scan state (agg[w]) from the entry counts of its quotients;
// wold[w] = its old incoming location (1 + the slot of the d_w-th run end, 0
// when d_w = 0).  Thread t: quotients [16 t, 16 t + 16), image slots
// [32 t, 32 t + 32).  The run ends at or after the window's first slot are
// ranked (s_rend[i] = slot of the (d_w + i)-th, i = 0 .. O), so the j-th
// occupied quotient p holds the slots [max(p, s_rend[j-1] + 1), s_rend[j]]
// (the decode rule of k_rd_emit); the batch adds its quotient counts.
// wp: exclusive occupied / run-end prefixes.
template <typename K>
__global__ __launch_bounds__(WS_THREADS) void k_wm_sum(RsqfDev f, const uint32_t* __restrict__ wp,
                                                       const K* __restrict__ B, const uint32_t* __restrict__ bstart,
                                                       uint4* __restrict__ agg, uint32_t* __restrict__ wold,
                                                       uint32_t* __restrict__ ctr) {
  __shared__ __align__(16) uint32_t s_cnt[RD_SLOTS / 2];  // batch entries per quotient, 16 bits each
  __shared__ int16_t s_rend[RD_SLOTS + 1];   // image slot of the (d + i)-th run end; -1: before the window
  __shared__ uint32_t s_tmp[WS_THREADS / 64];
  __shared__ uint4 s_wave[WS_THREADS / 64 + 1];
  const uint32_t w = blockIdx.x, t = threadIdx.x;
  const uint64_t bw = (uint64_t)w * RD_BLOCKS;
  const uint64_t qb = bw + (t >> 2), eb = bw + (t >> 1);
  const uint32_t o16 = qb < f.nb ? (uint32_t)(f.occ((uint32_t)qb) >> (16 * (t & 3))) & 0xFFFFu : 0u;
  const uint32_t e32 = eb < f.nb ? (uint32_t)(f.ends((uint32_t)eb) >> (32 * (t & 1))) : 0u;
  const uint32_t b0 = bstart[w], b1 = bstart[w + 1];
  const uint32_t d = wp[2 * (size_t)w] - wp[2 * (size_t)w + 1];
  for (uint32_t i = t; i < RD_SLOTS / 2; i += WS_THREADS) s_cnt[i] = 0u;
  if (t == 0) s_rend[0] = -1;  // d = 0: the previous run ended before the window
  // occupied ranks (high half) and run-end ranks (low half): at most 4096 and 8192
  const uint32_t pk = block_excl_scan<WS_THREADS>(((uint32_t)__popc(o16) << 16) | (uint32_t)__popc(e32), s_tmp);
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < (int)(WS_THREADS / 64); ++i) tot += s_tmp[i];
  const uint32_t O = tot >> 16, E = tot & 0xFFFFu;
  {
    uint32_t k = pk & 0xFFFFu;  // run ends before this thread's first slot
    for (uint32_t x = e32; x; x &= x - 1u) {
      ++k;  // rank of this run end
      if (k - d <= O) s_rend[k - d] = (int16_t)(32 * t + (uint32_t)__builtin_ctz(x));  // k < d wraps past O
    }
  }
  if (b1 - b0 <= WM_CAP) {
    const K kbase = (K)((uint64_t)w * RD_SLOTS) << f.r;
    for (uint32_t i = b0 + t; i < b1; i += WS_THREADS) wm_cnt_add(s_cnt, (uint32_t)((B[i] - kbase) >> f.r), 1u);
  }
  __syncthreads();
  const bool bad = d + O > E;  // a run end the image does not hold
  RS x = rs_none();
  uint32_t j = pk >> 16;  // occupied quotients before this thread's first one
  // this thread's 16 batch counts in two 16-byte reads
  uint32_t bc[8];
  {
    const uint4* c4 = reinterpret_cast<const uint4*>(s_cnt) + 2 * t;
    const uint4 lo = c4[0], hi = c4[1];
    bc[0] = lo.x, bc[1] = lo.y, bc[2] = lo.z, bc[3] = lo.w, bc[4] = hi.x, bc[5] = hi.y, bc[6] = hi.z, bc[7] = hi.w;
  }
  int32_t pe = bad ? 0 : s_rend[j];  // end of the run before this thread's first occupied quotient
#pragma unroll
  for (uint32_t i = 0; i < 16; ++i) {
    const uint32_t p = 16 * t + i;
    uint32_t n = (bc[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
    if ((o16 >> i) & 1u) {
      ++j;
      if (!bad) {
        const int32_t e = s_rend[j];
        const int32_t st = (int32_t)p > pe + 1 ? (int32_t)p : pe + 1;
        if (e >= st) n += (uint32_t)(e - st + 1);
        pe = e;
      }
    }
    if (n) rs_push_run(x, w * RD_SLOTS + p, n);
  }
  RS ex;
  const RS all = rs_block_n<WS_THREADS / 64>(x, &ex, s_wave);
  if (t == 0) {
    agg[w] = rs_pack(all);
    wold[w] = d && !bad ? w * RD_SLOTS + (uint32_t)(s_rend[0] + 1) : 0u;
    if (bad || all.c > WM_CAP || b1 - b0 > WM_CAP) ctr[CTR_MISC] = 1u;
  }
}

// The handover bound of every window (wsave), from the scanned states (each
// window's exclusive prefix: lp = its new incoming location) and wold; more
// than WM_SIDE blocks to hand over raises the fallback flag.
__global__ __launch_bounds__(256) void k_wm_check(const uint4* __restrict__ pre, const uint32_t* __restrict__ wold,
                                                  const uint32_t* __restrict__ bstart, uint32_t W,
                                                  uint32_t* __restrict__ wsave, uint8_t* __restrict__ wclean,
                                                  uint32_t* __restrict__ ctr) {
  const uint32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= W) return;
  const uint32_t lp = rs_unpack(pre[x]).lp, o = wold[x];
  const uint64_t h = lp > o ? lp : o;
  if (h > (uint64_t)x * RD_SLOTS + 64 * WM_SIDE) ctr[CTR_MISC] = 1u;
  wsave[x] = (uint32_t)h;
  // no batch entry and the same incoming location: the canonical layout of
  // the window's slots is the one stored (its outgoing location is unchanged
  // too), so the window is not rewritten -- unless its runs reach into the
  // next window, whose cleared blocks would then lose them
  const uint32_t s0 = x * RD_SLOTS;
  bool clean = bstart[x + 1] == bstart[x] && (lp > s0 ? lp : s0) == (o > s0 ? o : s0);
  if (clean && x + 1 < W) {
    const uint32_t lp1 = rs_unpack(pre[x + 1]).lp, o1 = wold[x + 1];
    clean = (lp1 > o1 ? lp1 : o1) <= s0 + RD_SLOTS;
  }
  wclean[x] = clean ? 1 : 0;
}

// Blocks of window w below its handover bound: run ends and remainders to
// the side copy (WM_SIDE blocks per window), then cleared in the table.
__global__ __launch_bounds__(64) void k_wm_save(uint8_t* __restrict__ base, uint32_t nb, uint32_t r,
                                                const uint32_t* __restrict__ wsave,
                                                const uint8_t* __restrict__ wclean, uint64_t* __restrict__ side_end,
                                                uint32_t* __restrict__ side_rem, const uint32_t* __restrict__ ctr) {
  if (ctr[CTR_ERR] || ctr[CTR_MISC]) return;
  const uint32_t w = blockIdx.x, t = threadIdx.x;
  // a window that is not rewritten keeps its blocks: its neighbour reads the
  // copy and ORs in the same bits it finds there
  const bool keep = wclean[w] != 0;
  const uint64_t s0 = (uint64_t)w * RD_SLOTS, h = wsave[w];
  if (h <= s0) return;
  const uint64_t bw = (uint64_t)w * RD_BLOCKS;
  uint64_t nsv = (h - s0 + 63) / 64;  // <= WM_SIDE: k_wm_check
  if (nsv > WM_SIDE) nsv = WM_SIDE;
  if (bw + nsv > nb) nsv = nb - bw;
  uint64_t* meta = reinterpret_cast<uint64_t*>(base);
  uint32_t* remg = reinterpret_cast<uint32_t*>(base + rsqf_meta_bytes(nb));
  const uint32_t rw = 2 * r;
  for (uint32_t i = t; i < nsv; i += 64) {
    side_end[(uint64_t)w * WM_SIDE + i] = meta[4 * (bw + i) + 1];
    if (!keep) meta[4 * (bw + i) + 1] = 0ull;
  }
  for (uint32_t i = t; i < nsv * rw; i += 64) {
    side_rem[(uint64_t)w * WM_SIDE * rw + i] = remg[bw * rw + i];
    if (!keep) remg[bw * rw + i] = 0u;
  }
}

// Image-relative slot s's remainder in an image of remainder words (RW per block).
template <int R, uint32_t RW>
__device__ __forceinline__ uint32_t wm_get_rem(const uint32_t* s_rem, uint32_t s, uint32_t rr) {
  if (R == 8) return reinterpret_cast<const uint8_t*>(s_rem)[s];
  if (R == 16) return reinterpret_cast<const uint16_t*>(s_rem)[s];
  const uint32_t bit = (s & 63) * rr;
  const uint32_t* blk = s_rem + (s >> 6) * RW;
  const uint32_t lo = blk[bit >> 5] >> (bit & 31);
  const uint32_t hi = (bit & 31) + rr > 32 ? blk[(bit >> 5) + 1] << (32 - (bit & 31)) : 0u;
  return (lo | hi) & ((1u << rr) - 1u);
}
template <int R, uint32_t RW>
__device__ __forceinline__ void wm_put_rem(uint32_t* s_rem, uint32_t s, uint32_t rr, uint32_t v) {
  if (R == 8) {
    reinterpret_cast<uint8_t*>(s_rem)[s] = (uint8_t)v;
  } else if (R == 16) {
    reinterpret_cast<uint16_t*>(s_rem)[s] = (uint16_t)v;
  } else {
    const uint32_t bit = (s & 63) * rr;
    uint32_t* blk = s_rem + (s >> 6) * RW;
    atomicOr(&blk[bit >> 5], v << (bit & 31));
    if ((bit & 31) + rr > 32) atomicOr(&blk[(bit >> 5) + 1], v >> (32 - (bit & 31)));
  }
}

// Entries of a sorted LDS array at most v.
__device__ __forceinline__ uint32_t wm_rank_le(const uint32_t* a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Window w: decode E_w (saved blocks from the side copy), merge B_w, lay M_w
// out from the window's incoming state and write the window's blocks (plus
// ORs into the saved blocks of window w + 1 its entries reach).  Entries are
// held window-relative: (quotient - 4096 w) << r | remainder.  Every input
// load goes out at the start: table and side words are both read for the
// blocks that may be saved, and chosen once the bounds are in.
template <typename K, int R>
__global__ __launch_bounds__(WM_THREADS, (R == 8 ? 6 : 4)) void k_wm_write(uint8_t* __restrict__ base, uint32_t nb, uint32_t r,
                                                            uint32_t W, const uint32_t* __restrict__ wp,
                                                            const K* __restrict__ B,
                                                            const uint32_t* __restrict__ bstart,
                                                            const uint4* __restrict__ pre,
                                                            const uint32_t* __restrict__ wsave,
                                                            const uint8_t* __restrict__ wclean,
                                                            const uint64_t* __restrict__ side_end,
                                                            const uint32_t* __restrict__ side_rem,
                                                            const uint32_t* __restrict__ ctr) {
  if (ctr[CTR_ERR] || ctr[CTR_MISC] || wclean[blockIdx.x]) return;
  constexpr uint32_t RW = R ? 2 * (uint32_t)R : 32;  // remainder words per image block
  constexpr uint32_t RL = (WM_IMG * RW + WM_THREADS - 1) / WM_THREADS;
  __shared__ __align__(16) uint32_t s_rem[WM_IMG * RW];
  __shared__ __align__(16) uint32_t s_m[WM_CAP];  // E_w from the front, B_w's low words at the back
  __shared__ uint32_t s_u[RD_SLOTS / 2];          // occupied list; merge bitmap; neighbours' quotients
  __shared__ uint32_t s_end[WM_IMG * 2];
  __shared__ uint32_t s_occ[RD_BLOCKS * 2];
  __shared__ uint32_t s_tmp[WM_THREADS / 64];
  __shared__ uint4 s_wave[WM_WAVES + 1];
  const uint32_t w = blockIdx.x, t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const uint32_t rr = R ? (uint32_t)R : r, rw = 2 * rr;
  uint64_t* meta = reinterpret_cast<uint64_t*>(base);
  uint32_t* remg = reinterpret_cast<uint32_t*>(base + rsqf_meta_bytes(nb));
  const uint64_t bw = (uint64_t)w * RD_BLOCKS;
  const bool next = w + 1 < W;
  // ---- loads: bounds and states, then table and side words, then the batch
  const uint64_t h0 = wsave[w], h1 = next ? wsave[w + 1] : 0u;
  const uint32_t b0 = bstart[w], nB = bstart[w + 1] - b0;
  const uint32_t d = wp[2 * (size_t)w] - wp[2 * (size_t)w + 1];
  const RS X = rs_unpack(pre[w]);
  const uint32_t ib = t >> 2, sh = 16 * (t & 3);  // this thread's image slots [16 t, 16 t + 16): block ib
  uint64_t om = 0, et = 0, es = 0;
  if (bw + ib < nb) {
    if (ib < RD_BLOCKS) {
      om = meta[4 * (bw + ib)];
      et = meta[4 * (bw + ib) + 1];
      if (ib < WM_SIDE) es = side_end[(uint64_t)w * WM_SIDE + ib];
    } else if (next && ib - RD_BLOCKS < WM_SIDE) {
      es = side_end[(uint64_t)(w + 1) * WM_SIDE + (ib - RD_BLOCKS)];
    }
  }
  uint32_t tv[RL], sv[RL];
#pragma unroll
  for (uint32_t k = 0; k < RL; ++k) {
    const uint32_t i = t + k * WM_THREADS, j = i / rw;
    tv[k] = 0u;
    sv[k] = 0u;
    if (i < WM_IMG * rw && bw + j < nb) {
      if (j < RD_BLOCKS) {
        tv[k] = remg[bw * rw + i];
        if (j < WM_SIDE) sv[k] = side_rem[(uint64_t)w * WM_SIDE * rw + i];
      } else if (next && j - RD_BLOCKS < WM_SIDE) {
        sv[k] = side_rem[(uint64_t)(w + 1) * WM_SIDE * rw + (i - RD_BLOCKS * rw)];
      }
    }
  }
  // the batch's low words straight to LDS (the difference to the window base fits them)
  const uint32_t nBc = nB < WM_CAP ? nB : WM_CAP;
  {
    const uint32_t* B32 = reinterpret_cast<const uint32_t*>(B);
    constexpr uint32_t KW = sizeof(K) / 4;
    for (uint32_t k = 0; k * WM_THREADS < nBc; ++k) {
      const uint32_t bs = (k * WM_WAVES + wid) * 64;
      if (bs + lane < nBc)
        __builtin_amdgcn_global_load_lds(B32 + KW * ((uint64_t)b0 + bs + lane), s_m + (WM_CAP - nBc) + bs, 4, 0, 0);
    }
  }
  auto saved = [&](uint32_t i) { return 64 * (bw + i) < (i < RD_BLOCKS ? h0 : h1); };
#pragma unroll
  for (uint32_t k = 0; k < RL; ++k) {
    const uint32_t i = t + k * WM_THREADS;
    if (i < WM_IMG * rw) {
      const uint32_t j = i / rw;
      s_rem[j * RW + (i - j * rw)] = saved(j) ? sv[k] : tv[k];
    }
  }
  const uint32_t o16 = ib < RD_BLOCKS && t < 256 ? (uint32_t)(om >> sh) & 0xFFFFu : 0u;
  const uint32_t e16 = (uint32_t)((saved(ib) ? es : et) >> sh) & 0xFFFFu;
  // ---- decode: occupied ranks and run-end ranks, the occupied list, the entries
  const uint32_t pk = block_excl_scan<WM_THREADS>(((uint32_t)__popc(o16) << 16) | (uint32_t)__popc(e16), s_tmp);
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < (int)(WM_THREADS / 64); ++i) tot += s_tmp[i];
  const uint32_t O = tot >> 16;
  uint16_t* s_list = reinterpret_cast<uint16_t*>(s_u);
  {
    uint32_t g = pk >> 16;
    for (uint32_t x = o16; x; x &= x - 1u) s_list[g++] = (uint16_t)(16 * t + (uint32_t)__builtin_ctz(x));
  }
  __syncthreads();
  uint32_t qv[16];
  uint32_t m = 0;
  if ((pk & 0xFFFFu) < d + O) {  // else every slot of this thread lies after the window's last run
    uint32_t k = pk & 0xFFFFu;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) {
      const uint32_t s = 16 * t + i;
      const uint32_t Q = k - d < O ? s_list[k - d] : 0xFFFFu;  // k < d wraps past O
      qv[i] = Q;
      if (Q <= s) m |= 1u << i;
      k += (e16 >> i) & 1u;
    }
  }
  const uint32_t posE = block_excl_scan<WM_THREADS>((uint32_t)__popc(m), s_tmp);
  uint32_t nE = 0;
#pragma unroll
  for (int i = 0; i < (int)(WM_THREADS / 64); ++i) nE += s_tmp[i];
  if (nE + nB > WM_CAP) {  // unreachable: k_wm_sum counted the same entries and raised the flag
    rb_stage_wait();
    return;
  }
  const uint32_t nM = nE + nB;
  {
    uint32_t p = posE;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i)
      if ((m >> i) & 1u) s_m[p++] = (qv[i] << rr) | wm_get_rem<R, RW>(s_rem, 16 * t + i, rr);
  }
  // ---- stable merge: each batch entry marks its output position in a bitmap
  const uint32_t kb32 = (uint32_t)((uint64_t)w * RD_SLOTS << rr);
  const uint32_t* sB = s_m + (WM_CAP - nB);
  uint32_t* s_bmp = s_u;
  uint32_t* s_bpre = s_u + WM_CAP / 32;
  rb_stage_wait();  // the batch words are in LDS
  __syncthreads();  // every entry is written, the occupied list is no longer read
  uint32_t v[WM_ITEMS];
  const uint32_t i0 = WM_ITEMS * t;
  if (nB) {
    if (t < WM_CAP / 32) s_bmp[t] = 0u;
    __syncthreads();
    for (uint32_t j = t; j < nB; j += WM_THREADS) {
      const uint32_t p = j + wm_rank_le(s_m, nE, sB[j] - kb32);
      atomicOr(&s_bmp[p >> 5], 1u << (p & 31));
    }
    __syncthreads();
    static_assert(WM_CAP / 32 <= 256, "one wave scans the bitmap, 4 words per lane");
    if (t < 64) {
      uint32_t c[4], sum = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        c[i] = 4 * t + i < WM_CAP / 32 ? (uint32_t)__popc(s_bmp[4 * t + i]) : 0u;
        sum += c[i];
      }
      uint32_t run = wave_incl_add(sum) - sum;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (4 * t + i < WM_CAP / 32) s_bpre[4 * t + i] = run;
        run += c[i];
      }
    }
    __syncthreads();
    const uint32_t wd = s_bmp[i0 >> 5];
    uint32_t bi = s_bpre[i0 >> 5] + (uint32_t)__popc(wd & ((1u << (i0 & 31)) - 1u));
    uint32_t ai = i0 - bi;
    const uint32_t bits = wd >> (i0 & 31);
    if (i0 < nM) {
      // branch-free: one LDS read per entry at the chosen index
#pragma unroll
      for (uint32_t j = 0; j < WM_ITEMS; ++j) {
        const uint32_t isb = (bits >> j) & 1u;
        const uint32_t at = isb ? (WM_CAP - nB) + bi : ai;
        v[j] = i0 + j < nM ? s_m[at < WM_CAP ? at : 0u] - (isb ? kb32 : 0u) : 0u;
        bi += isb;
        ai += isb ^ 1u;
      }
    } else {
#pragma unroll
      for (uint32_t j = 0; j < WM_ITEMS; ++j) v[j] = 0u;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < WM_ITEMS; ++j) v[j] = i0 + j < nM ? s_m[i0 + j] : 0u;
  }
  const uint32_t cnt = nM > i0 ? (nM - i0 < WM_ITEMS ? nM - i0 : WM_ITEMS) : 0u;
  // ---- locations: this thread's entries on top of the window's incoming state
  RS a = rs_none();
#pragma unroll
  for (uint32_t j = 0; j < WM_ITEMS; ++j)
    if (j < cnt) rs_push(a, w * RD_SLOTS + (v[j] >> rr));
  __syncthreads();  // the bitmap is no longer read: the first quotients take its place
  uint16_t* s_q0 = reinterpret_cast<uint16_t*>(s_u);  // each thread's first quotient
  if (cnt) s_q0[t] = (uint16_t)(v[0] >> rr);
  for (uint32_t i = t; i < WM_IMG * RW; i += WM_THREADS) s_rem[i] = 0u;
  if (t < WM_IMG * 2) s_end[t] = 0u;
  // occupieds after the insert: the stored ones, and the batch's quotients (below)
  if (t < 4 * RD_BLOCKS) reinterpret_cast<uint16_t*>(s_occ)[t] = (uint16_t)o16;
  RS ex;
  (void)rs_block_n<WM_WAVES>(a, &ex, s_wave);  // two barriers: the cleared images and the neighbours are visible after it
  const RS me = rs_combine(X, ex);
  for (uint32_t j = t; j < nB; j += WM_THREADS) {
    const uint32_t q = (sB[j] - kb32) >> rr;
    atomicOr(&s_occ[q >> 5], 1u << (q & 31));
  }
  const uint32_t s0 = w * RD_SLOTS;
  const uint32_t rmask = (1u << rr) - 1u;
  constexpr uint32_t NONE = RD_SLOTS + 64;  // past every window-relative quotient
  // quotient of the entry after this thread's last one
  const uint32_t nq_last = i0 + cnt < nM ? s_q0[t + 1] : NONE;
  // every entry at its location: remainder, and the run end where the next
  // entry's quotient differs (whole threads: no count checks)
  auto place = [&](auto full) {
    constexpr bool FULL = decltype(full)::value;
    int32_t prev = (int32_t)me.lp - 1;
#pragma unroll
    for (uint32_t j = 0; j < WM_ITEMS; ++j) {
      if (FULL || j < cnt) {
        const uint32_t fqr = v[j] >> rr;
        const int32_t fq = (int32_t)(s0 + fqr);
        const int32_t lc = fq > prev + 1 ? fq : prev + 1;
        const uint32_t rel = (uint32_t)lc - s0;  // < 4096 + 64 WM_SIDE: k_wm_check
        const uint32_t nq = j + 1 < WM_ITEMS && (FULL || j + 1 < cnt) ? v[j + 1] >> rr : nq_last;
        if (rel < WM_IMG * 64) {
          wm_put_rem<R, RW>(s_rem, rel, rr, v[j] & rmask);
          if (nq != fqr) atomicOr(&s_end[rel >> 5], 1u << (rel & 31));
        }
        prev = lc;
      }
    }
  };
  if (cnt == WM_ITEMS) place(std::true_type{});
  else if (cnt) place(std::false_type{});
  __syncthreads();
  // ---- block offsets (one wave): the last entry before block b's first
  // quotient ends the run of the window's k-th occupied quotient, k = its
  // occupieds before the block -- the k-th run end of the window in s_end --
  // or, for k = 0, is the last entry of earlier windows
  if (t < RD_BLOCKS) {
    static_assert(WM_IMG * 2 == 4 * RD_BLOCKS, "four run-end words per lane");
    uint32_t es = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) es += (uint32_t)__popc(s_end[4 * t + i]);
    const uint32_t ep = wave_incl_add(es) - es;  // run ends before word 4 t
    const uint32_t oc = (uint32_t)__popc(s_occ[2 * t]) + (uint32_t)__popc(s_occ[2 * t + 1]);
    const uint32_t k = wave_incl_add(oc) - oc;
    uint32_t lo = 0, hi = RD_BLOCKS;  // the last word group whose run ends before it are fewer than k
#pragma unroll
    for (int it = 0; it < 6; ++it) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t em = (uint32_t)__shfl((int)ep, (int)mid);
      if (em < k) lo = mid;
      else hi = mid;
    }
    uint32_t rest = k - (uint32_t)__shfl((int)ep, (int)lo), word = 4 * lo + 3;
    bool found = false;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t c = (uint32_t)__popc(s_end[4 * lo + i]);
      if (!found && rest <= c) {
        word = 4 * lo + i;
        found = true;
      } else if (!found) {
        rest -= c;
      }
    }
    const int64_t L = k ? (int64_t)s0 + 32 * word + select_bit64(s_end[word], rest) : (int64_t)X.lp - 1;
    if (bw + t < nb) {
      const int64_t sb = 64 * (int64_t)(bw + t);
      meta[4 * (bw + t) + 2] = L >= sb ? (uint64_t)((L - sb + 1) & 0xFF) : 0ull;
    }
  }
  // ---- stores: whole words where the window owns the block, ORs into saved blocks
  if (t < RD_BLOCKS && bw + t < nb) meta[4 * (bw + t)] = ((uint64_t)s_occ[2 * t + 1] << 32) | s_occ[2 * t];
  if (t < WM_IMG && bw + t < nb) {
    const uint64_t b = bw + t;
    const uint64_t e = ((uint64_t)s_end[2 * t + 1] << 32) | s_end[2 * t];
    if (saved(t)) {
      if (e) atomicOr((unsigned long long*)&meta[4 * b + 1], (unsigned long long)e);
    } else if (t < RD_BLOCKS) {
      meta[4 * b + 1] = e;
    }
  }
  for (uint32_t i = t; i < WM_IMG * rw; i += WM_THREADS) {
    const uint32_t j = i / rw;
    const uint32_t x = s_rem[j * RW + (i - j * rw)];
    if (bw + j < nb) {
      if (saved(j)) {
        if (x) atomicOr(&remg[bw * rw + i], x);
      } else if (j < RD_BLOCKS) {
        remg[bw * rw + i] = x;
      }
    }
  }
}

// ---------------------------------------------------------------- pipelines
struct RsqfGeom {
  uint32_t q, r, nb, bits;
  uint64_t cap;
};
static RsqfGeom rgeom(const gqf_rsqf& f) {
  RsqfGeom g;
  g.q = f.qbits;
  g.r = f.rbits;
  g.nb = f.numBlocks;
  g.bits = f.qbits + f.rbits;
  g.cap = (uint64_t)g.nb * 64;
  return g;
}
static RsqfDev rdev(const gqf_rsqf& f) {
  RsqfDev d;
  d.meta = reinterpret_cast<const uint64_t*>(f.blocks);
  d.rem = reinterpret_cast<const uint8_t*>(f.blocks) + rsqf_meta_bytes(f.numBlocks);
  d.nb = f.numBlocks;
  d.r = f.rbits;
  return d;
}

static inline uint32_t rd_windows(uint32_t nb) { return (nb + RD_BLOCKS - 1) / RD_BLOCKS; }

// Decoded copies of 64-bit fingerprints of at most 34 bits are packed (low
// words plus the three indices where the high word steps up; FpArray).
static inline bool rsqf_packed(const RsqfGeom& g, size_t kb) { return kb == 8 && g.bits <= 34; }

static size_t rsqf_decode_ws(const RsqfGeom& g, size_t kb) {
  const uint32_t W = rd_windows(g.nb);
  const size_t eb = rsqf_packed(g, kb) ? 4 : kb;
  return aligned_bytes((size_t)W * 8) + aligned_bytes((size_t)W * 4) + 256 + aligned_bytes((g.cap + 1) * eb) + 512 +
         scan_rows_bytes(W, 2) + scan_rows_bytes(W, 1);
}

// The stored fingerprints in ascending order; their count is left in device
// memory (*m_dev).  *hb: the bounds of the packed form (rsqf_packed), else
// NULL.
template <typename K>
static gqf_status rsqf_decode(hipStream_t s, Arena& ar, const gqf_rsqf& f, K** out, const uint32_t** m_dev,
                              const uint32_t** hb_out) {
  const RsqfDev d = rdev(f);
  const RsqfGeom g = rgeom(f);
  const bool pk = rsqf_packed(g, sizeof(K));
  const uint32_t nhv = pk ? (1u << (g.bits - 32)) - 1u : 0u;
  const uint32_t W = rd_windows(f.numBlocks);
  uint32_t* wp = ar.take<uint32_t>((size_t)W * 2);
  uint32_t* wc = ar.take<uint32_t>(W);
  uint32_t* tot = ar.take<uint32_t>(4);
  uint32_t* hb = pk ? ar.take<uint32_t>(4) : nullptr;
  if (ar.used > ar.cap) {
    set_error("internal: decode workspace overflow");
    return GQF_ERR_INVALID;
  }
  const double meta = (double)f.numBlocks * 16;  // occupieds + run ends
  GQF_LAUNCH_B(meta, k_rd_pop, dim3((W + 3) / 4), dim3(256), 0, s, d.meta, f.numBlocks, W, wp, hb, nhv);
  GQF_TRY(scan_rows<2>(s, ar, wp, W, tot));
  GQF_LAUNCH_B(meta, k_rd_count, dim3(W), dim3(RD_THREADS), 0, s, d, wp, wc);
  GQF_TRY(scan_rows<1>(s, ar, wc, W, tot + 2));
  // the count stays on the device (tot[2]): later kernels read it, the host
  // plans for the capacity
  *m_dev = tot + 2;
  // packed: 4 bytes per entry (K* typed, holding 32-bit words)
  K* fps = pk ? reinterpret_cast<K*>(ar.take<uint32_t>((size_t)f.numBlocks * 64 + 1))
              : ar.take<K>((size_t)f.numBlocks * 64 + 1);
  if (ar.used > ar.cap) {
    set_error("internal: decode workspace overflow");
    return GQF_ERR_INVALID;
  }
  // table read (remainders + run ends) + the entries written (count on the device)
  const double bytes = (double)rsqf_dev_bytes(f.qbits, f.rbits) - (double)f.numBlocks * 16;
  const double per = pk ? 4.0 : (double)sizeof(K);
#define GQF_RD_EMIT(R_, PK_) \
  GQF_LAUNCH_BC(bytes, *m_dev, per, (k_rd_emit<K, R_, PK_>), dim3(W), dim3(RD_THREADS), 0, s, d, wp, wc, fps, hb, nhv)
  if (pk) {
    if (f.rbits == 8) GQF_RD_EMIT(8, true);
    else if (f.rbits == 16) GQF_RD_EMIT(16, true);
    else GQF_RD_EMIT(0, true);
  } else {
    if (f.rbits == 8) GQF_RD_EMIT(8, false);
    else if (f.rbits == 16) GQF_RD_EMIT(16, false);
    else GQF_RD_EMIT(0, false);
  }
#undef GQF_RD_EMIT
  *out = fps;
  *hb_out = hb;
  return GQF_OK;
}

static size_t rsqf_rebuild_ws(uint64_t N) {
  const RbPlan p = rb_plan(N ? N : 1);
  return aligned_bytes((size_t)(p.nsub + 1) * 4) + aligned_bytes((size_t)p.G * 16) + 256 + 1024;
}

// Canonical layout of A (MERGE: merged with B; KEEP_FLAGS: the flagged
// entries of A) written over the table.
// na_dev (may be NULL): A's count on the device; na is then an upper bound
// that sizes the grids.  PK: A is a packed decode with its bounds at hb.
template <typename K, int MODE, bool MERGE, bool PK>
static gqf_status rsqf_rebuild_t(hipStream_t s, DeviceState& ds, Arena& ar, const gqf_rsqf& f, const K* A,
                                 uint32_t na, const uint32_t* na_dev, const K* B, uint32_t nbk,
                                 const uint8_t* flags, const uint32_t* hb) {
  const RsqfGeom g = rgeom(f);
  const uint64_t N = (uint64_t)na + (MERGE ? nbk : 0u);
  const size_t tbytes = rsqf_dev_bytes(g.q, g.r);
  if (N == 0) {
    GQF_LAUNCH(k_reset_counters, dim3(1), dim3(64), 0, s, ds.counters, (int)CTR_OP, (int)CTR_TOTAL);
    GQF_LAUNCH_B((double)tbytes, k_rr_clear, dim3(1024), dim3(256), 0, s, reinterpret_cast<uint8_t*>(f.blocks),
                 (uint64_t)tbytes, ds.counters);
    return GQF_OK;
  }
  const RbPlan p = rb_plan(N, MERGE ? RR_MERGE_PER : 1u);
  uint32_t* splits = ar.take<uint32_t>(p.nsub + 1);
  uint4* agg = ar.take<uint4>(p.G);
  uint4* total = ar.take<uint4>(1);
  if (ar.used > ar.cap) {
    set_error("internal: rebuild workspace overflow");
    return GQF_ERR_INVALID;
  }
  // bytes read per stored entry (the host-known entries are new ones, full width)
  const double per = (double)((PK ? 4 : sizeof(K)) + (MODE == KEEP_FLAGS ? 1 : 0));
  const double nknown = na_dev ? (double)(MERGE ? nbk : 0u) : (double)N;  // entries the host knows of
  const double kper = na_dev ? (double)sizeof(K) : per;
  if (MERGE)
    GQF_LAUNCH((k_rb_splits<K, PK>), dim3(((p.nsub + 1) * RB_SPLIT_LANES + 255) / 256), dim3(256), 0, s, A, na, na_dev, B, nbk, p.nsub,
               splits, hb);
  GQF_LAUNCH_BC(nknown * kper, na_dev, per, (k_rr_agg<K, MODE, MERGE, PK>), dim3(p.G), dim3(RB_THREADS), 0, s, A,
                na, na_dev, B, nbk, flags, splits, g.q, g.r, p.chunk, agg, hb, ds.counters);
  GQF_TRY(rb_scan(s, ar, agg, p.G, g.cap, ds.counters, total));
  GQF_LAUNCH_B((double)tbytes, k_rr_clear, dim3(1024), dim3(256), 0, s, reinterpret_cast<uint8_t*>(f.blocks),
               (uint64_t)tbytes, ds.counters);
  const double wbytes = nknown * kper + (double)tbytes;
#define GQF_RR_WRITE(R_)                                                                                          \
  GQF_LAUNCH_BC(wbytes, na_dev, per, (k_rr_write<K, MODE, MERGE, R_, PK>), dim3(p.G), dim3(RB_THREADS), 0, s,    \
                reinterpret_cast<uint8_t*>(f.blocks), g.nb, g.q, g.r, A, na, na_dev, B, nbk, flags, splits, p.chunk, \
                agg, ds.counters, hb)
  if (g.r == 8) GQF_RR_WRITE(8);
  else if (g.r == 16) GQF_RR_WRITE(16);
  else GQF_RR_WRITE(0);
#undef GQF_RR_WRITE
  GQF_LAUNCH(k_rr_tail, dim3(1), dim3(64), 0, s, reinterpret_cast<uint8_t*>(f.blocks), total, ds.counters);
  return GQF_OK;
}
template <typename K, int MODE, bool MERGE>
static gqf_status rsqf_rebuild(hipStream_t s, DeviceState& ds, Arena& ar, const gqf_rsqf& f, const K* A,
                               uint32_t na, const uint32_t* na_dev, const K* B, uint32_t nbk,
                               const uint8_t* flags, const uint32_t* hb = nullptr) {
  if constexpr (sizeof(K) == 8)
    if (hb) return rsqf_rebuild_t<K, MODE, MERGE, true>(s, ds, ar, f, A, na, na_dev, B, nbk, flags, hb);
  return rsqf_rebuild_t<K, MODE, MERGE, false>(s, ds, ar, f, A, na, na_dev, B, nbk, flags, nullptr);
}

static gqf_status rsqf_args(const gqf_rsqf& f, int n, const uint32_t* keys) {
  GQF_TRY(rsqf_check(f));
  if (n < 0 || (uint64_t)n >= GQF_MAX_ITEMS || (n > 0 && !keys)) {
    set_error("invalid key array (n=%d)", n);
    return GQF_ERR_INVALID;
  }
  return GQF_OK;
}

template <typename K>
static gqf_status rsqf_build_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(n, sizeof(K), false) + rsqf_rebuild_ws(n)));
  Arena ar(ds.ws, ds.ws_cap);
  K* sorted;
  GQF_TRY(radix_sort<K>(cs.stream, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &sorted, nullptr, cs.ds));
  return rsqf_rebuild<K, KEEP_ALL, false>(cs.stream, ds, ar, f, sorted, n, nullptr, (const K*)nullptr, 0u, nullptr);
}

template <typename K>
static size_t rsqf_insert_ws(const RsqfGeom& g, uint32_t n) {
  return rsqf_decode_ws(g, sizeof(K)) + sort_workspace_bytes(n, sizeof(K), false) + rsqf_rebuild_ws(g.cap + n);
}

template <typename K>
static gqf_status rsqf_insert_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, rsqf_insert_ws<K>(g, n)));
  Arena ar(ds.ws, ds.ws_cap);
  K* A;
  const uint32_t* m_dev = nullptr;
  const uint32_t* hb = nullptr;
  GQF_TRY(rsqf_decode<K>(cs.stream, ar, f, &A, &m_dev, &hb));
  K* B;
  GQF_TRY(radix_sort<K>(cs.stream, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &B, nullptr, cs.ds));
  // at most cap stored entries: the grids are sized for cap + n
  return rsqf_rebuild<K, KEEP_ALL, true>(cs.stream, ds, ar, f, A, (uint32_t)g.cap, m_dev, B, n, nullptr, hb);
}

// Workspace of the windowed insert (beside the batch's sort).
static size_t rsqf_win_ws(const RsqfGeom& g) {
  const uint32_t W = rd_windows(g.nb);
  return aligned_bytes((size_t)W * 8) + 256 + scan_rows_bytes(W, 2) + aligned_bytes(((size_t)W + 1) * 4) +
         aligned_bytes((size_t)W * 16) + 2 * aligned_bytes((size_t)W * 4) + aligned_bytes(W) +
         aligned_bytes((size_t)W * WM_SIDE * 8) +
         aligned_bytes((size_t)W * WM_SIDE * 8 * g.r) + aligned_bytes(((W + RBS_PER - 1) / RBS_PER) * 16) + 1024;
}
// Tables the windowed insert takes: one scan state per window.
static bool rsqf_win_fits(const RsqfGeom& g) { return rd_windows(g.nb) <= RB_MAX_CHUNKS; }

// The windowed insert of the sorted fingerprints B (see "windowed insert").
// Leaves ctr[CTR_MISC] != 0 when a window needs the rebuild instead; the
// table is then untouched.
template <typename K>
static gqf_status rsqf_insert_win(hipStream_t s, DeviceState& ds, Arena& ar, const gqf_rsqf& f, const K* B,
                                  uint32_t n) {
  const RsqfGeom g = rgeom(f);
  const RsqfDev d = rdev(f);
  const uint32_t W = rd_windows(g.nb);
  uint32_t* wp = ar.take<uint32_t>((size_t)W * 2);
  uint32_t* tot = ar.take<uint32_t>(4);
  uint32_t* bstart = ar.take<uint32_t>((size_t)W + 1);
  uint4* agg = ar.take<uint4>(W);
  uint32_t* wold = ar.take<uint32_t>(W);
  uint32_t* wsave = ar.take<uint32_t>(W);
  uint8_t* wclean = ar.take<uint8_t>(W);
  uint64_t* side_end = ar.take<uint64_t>((size_t)W * WM_SIDE);
  uint32_t* side_rem = ar.take<uint32_t>((size_t)W * WM_SIDE * 2 * g.r);
  if (ar.used > ar.cap) {
    set_error("internal: windowed insert workspace overflow");
    return GQF_ERR_INVALID;
  }
  const double meta = (double)g.nb * 16;  // occupieds + run ends
  GQF_LAUNCH_B(meta, k_rd_pop, dim3((W + 3) / 4), dim3(256), 0, s, d.meta, g.nb, W, wp, (uint32_t*)nullptr, 0u);
  GQF_TRY(scan_rows<2>(s, ar, wp, W, tot));
  GQF_LAUNCH((k_wm_bounds<K>), dim3(((W + 1) * WB_LANES + 255) / 256), dim3(256), 0, s, B, n, g.r, W, bstart,
             ds.counters);
  GQF_LAUNCH_B(meta * 1.5 + (double)n * sizeof(K), (k_wm_sum<K>), dim3(W), dim3(WS_THREADS), 0, s, d, wp, B, bstart,
               agg, wold, ds.counters);
  GQF_TRY(rb_scan(s, ar, agg, W, g.cap, ds.counters, nullptr));
  GQF_LAUNCH(k_wm_check, dim3((W + 255) / 256), dim3(256), 0, s, agg, wold, bstart, W, wsave, wclean, ds.counters);
  GQF_LAUNCH(k_wm_save, dim3(W), dim3(64), 0, s, reinterpret_cast<uint8_t*>(f.blocks), g.nb, g.r, wsave, wclean,
             side_end, side_rem, ds.counters);
  // the table read once and written once (plus the batch)
  const double wbytes = 2.0 * (double)rsqf_dev_bytes(g.q, g.r) + (double)n * sizeof(K);
#define GQF_WM_WRITE(R_)                                                                                          \
  GQF_LAUNCH_B(wbytes, (k_wm_write<K, R_>), dim3(W), dim3(WM_THREADS), 0, s, reinterpret_cast<uint8_t*>(f.blocks), \
               g.nb, g.r, W, wp, B, bstart, agg, wsave, wclean, side_end, side_rem, ds.counters)
  if (g.r == 8) GQF_WM_WRITE(8);
  else if (g.r == 16) GQF_WM_WRITE(16);
  else GQF_WM_WRITE(0);
#undef GQF_WM_WRITE
  return GQF_OK;
}

template <typename K>
static gqf_status rsqf_insert_win_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(n, sizeof(K), false) + rsqf_win_ws(g)));
  Arena ar(ds.ws, ds.ws_cap);
  K* B;
  GQF_TRY(radix_sort<K>(cs.stream, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &B, nullptr, cs.ds));
  return rsqf_insert_win<K>(cs.stream, ds, ar, f, B, n);
}

template <typename K>
static size_t rsqf_delete_ws(const RsqfGeom& g, uint32_t n) {
  return rsqf_decode_ws(g, sizeof(K)) + sort_workspace_bytes(n, sizeof(K), false) + aligned_bytes(g.cap + 1) +
         rsqf_rebuild_ws(g.cap);
}

template <typename K>
static gqf_status rsqf_delete_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, rsqf_delete_ws<K>(g, n)));
  Arena ar(ds.ws, ds.ws_cap);
  K* A;
  const uint32_t* m_dev = nullptr;
  const uint32_t* hb = nullptr;
  GQF_TRY(rsqf_decode<K>(cs.stream, ar, f, &A, &m_dev, &hb));
  K* D;
  GQF_TRY(radix_sort<K>(cs.stream, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &D, nullptr, cs.ds));
  const uint32_t mcap = (uint32_t)g.cap;  // the grids cover the capacity; kernels read the count
  uint8_t* keep = ar.take<uint8_t>((size_t)mcap + 1);
  if (ar.used > ar.cap) {
    set_error("internal: delete workspace overflow");
    return GQF_ERR_INVALID;
  }
  const double dper = (double)((hb ? 4 : sizeof(K)) + 1);
  if constexpr (sizeof(K) == 8) {
    if (hb)
      GQF_LAUNCH_BC(0.0, m_dev, dper, (k_delete_mark<K, true>), dim3((mcap + DM_TILE - 1) / DM_TILE), dim3(DM_THREADS),
                    0, cs.stream, A, mcap, m_dev, D, n, keep, hb);
  }
  if (!hb)
    GQF_LAUNCH_BC(0.0, m_dev, dper, (k_delete_mark<K, false>), dim3((mcap + DM_TILE - 1) / DM_TILE), dim3(DM_THREADS),
                  0, cs.stream, A, mcap, m_dev, D, n, keep, (const uint32_t*)nullptr);
  return rsqf_rebuild<K, KEEP_FLAGS, false>(cs.stream, ds, ar, f, A, mcap, m_dev, (const K*)nullptr, 0u, keep, hb);
}

// Keys left out of a partial insert: fingerprint >= cut -> -1.
__global__ __launch_bounds__(256) void k_rk_mark(const uint32_t* __restrict__ keys, uint32_t n, uint32_t bits,
                                                 uint64_t cut, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n && fingerprint(keys[i], bits) >= cut) out[i] = RS_NOT_FOUND;
}

static void launch_rlookup(hipStream_t s, const gqf_rsqf& f, uint32_t n, const uint32_t* keys, uint32_t* out) {
  const RsqfDev d = rdev(f);
  const dim3 grid((n + 255) / 256);
  if (f.rbits == 8) GQF_LAUNCH_B((double)n * (8 + 128), k_rsqf_lookup<8>, grid, dim3(256), 0, s, d, f.qbits, keys, n, out);
  else GQF_LAUNCH_B((double)n * (8 + 128), k_rsqf_lookup<0>, grid, dim3(256), 0, s, d, f.qbits, keys, n, out);
}

}  // namespace gqf

using namespace gqf;

extern "C" {

uint32_t gqf_rsqf_num_blocks(unsigned int q) { return rsqf_blocks(q); }
size_t gqf_rsqf_table_bytes(unsigned int q, unsigned int r) { return rsqf_dev_bytes(q, r); }
size_t gqf_rsqf_ref_bytes(unsigned int q, unsigned int r) { return (size_t)rsqf_blocks(q) * rsqf_ref_block_bytes(r); }

gqf_status gqf_rsqf_init(gqf_rsqf* cqf, unsigned int q, unsigned int r) {
  if (!cqf) return GQF_ERR_INVALID;
  gqf_rsqf t;
  t.qbits = q;
  t.rbits = r;
  t.numBlocks = rsqf_blocks(q);
  t.blocks = reinterpret_cast<void*>(1);
  GQF_TRY(rsqf_check(t));
  const size_t bytes = rsqf_dev_bytes(q, r);
  void* p = nullptr;
  GQF_HIP(hipMalloc(&p, bytes));
  GQF_HIP(hipMemset(p, 0, bytes));
  GQF_HIP(hipDeviceSynchronize());
  t.blocks = p;
  *cqf = t;
  return GQF_OK;
}

gqf_status gqf_rsqf_destroy(gqf_rsqf* cqf) {
  if (!cqf) return GQF_ERR_INVALID;
  if (cqf->blocks) GQF_HIP(hipFree(cqf->blocks));
  cqf->blocks = nullptr;
  return GQF_OK;
}

gqf_status gqf_rsqf_bulk_build(gqf_rsqf cqf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms) {
  GQF_TRY(rsqf_args(cqf, n, d_keys));
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  if (cqf.qbits + cqf.rbits <= 32) GQF_TRY(rsqf_build_t<uint32_t>(cqf, (uint32_t)n, d_keys, cs));
  else GQF_TRY(rsqf_build_t<uint64_t>(cqf, (uint32_t)n, d_keys, cs));
  return cs.end(ms, true);
}

// insertGPU as gqf_rsqf_insert; *lp_end (may be NULL) receives 1 + the last
// location of the merged layout (saturating), also when it does not fit.
static gqf_status rsqf_insert_call(gqf_rsqf cqf, int n, const uint32_t* d_keys, int32_t* d_out, gqf_stream stream,
                                   float* ms, uint32_t* lp_end) {
  GQF_TRY(rsqf_args(cqf, n, d_keys));
  const bool wide = cqf.qbits + cqf.rbits > 32;
  float ms_win = 0.f;
  if (n > 0 && rsqf_insert_path() != GQF_INSERT_REBUILD && rsqf_win_fits(rgeom(cqf))) {
    CallScope cs;
    GQF_TRY(cs.begin(stream));
    if (!wide) GQF_TRY(rsqf_insert_win_t<uint32_t>(cqf, (uint32_t)n, d_keys, cs));
    else GQF_TRY(rsqf_insert_win_t<uint64_t>(cqf, (uint32_t)n, d_keys, cs));
    if (d_out) launch_rlookup(cs.stream, cqf, (uint32_t)n, d_keys, reinterpret_cast<uint32_t*>(d_out));
    const gqf_status st = cs.end(&ms_win, true);
    if (st != GQF_OK && st != GQF_ERR_CAPACITY) return st;
    // a window the windowed path cannot take: nothing was written, rebuild
    if (cs.ds->h_pinned[CTR_MISC] == 0) {
      if (ms) *ms = ms_win;
      if (lp_end) *lp_end = cs.ds->h_pinned[CTR_MISC + 1];
      return st;
    }
  }
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  if (!wide) GQF_TRY(rsqf_insert_t<uint32_t>(cqf, (uint32_t)n, d_keys, cs));
  else GQF_TRY(rsqf_insert_t<uint64_t>(cqf, (uint32_t)n, d_keys, cs));
  if (d_out && n > 0) launch_rlookup(cs.stream, cqf, (uint32_t)n, d_keys, reinterpret_cast<uint32_t*>(d_out));
  const gqf_status st = cs.end(ms, true);
  if (ms) *ms += ms_win;
  if (lp_end) *lp_end = n > 0 ? cs.ds->h_pinned[CTR_MISC + 1] : 0u;
  return st;
}

gqf_status gqf_rsqf_insert(gqf_rsqf cqf, int n, const uint32_t* d_keys, int32_t* d_out, gqf_stream stream, float* ms) {
  return rsqf_insert_call(cqf, n, d_keys, d_out, stream, ms, nullptr);
}

gqf_status gqf_rsqf_insert_report(gqf_rsqf cqf, int n, const uint32_t* d_keys, int32_t* d_out, uint32_t* n_failed,
                                  gqf_stream stream, float* ms) {
  if (n_failed) *n_failed = 0;
  uint32_t lp = 0;
  float t0 = 0.f;
  const gqf_status st = rsqf_insert_call(cqf, n, d_keys, d_out, stream, &t0, &lp);
  if (ms) *ms = t0;
  if (st != GQF_ERR_CAPACITY) return st;
  // The batch does not fit (nothing was written).  With e = 1 + the last
  // location of the full merge, leaving out the k = e - S largest new
  // fingerprints moves every earlier entry back by k and leaves only stored
  // entries after them, so the rest fits; the cut is rounded down to whole
  // fingerprints (every copy of a fingerprint shares its fate).
  const RsqfGeom g = rgeom(cqf);
  const auto t_start = std::chrono::steady_clock::now();
  std::vector<uint32_t> hk((size_t)n);
  GQF_HIP(hipMemcpy(hk.data(), d_keys, (size_t)n * 4, hipMemcpyDeviceToHost));
  std::vector<uint64_t> fp((size_t)n);
  for (int i = 0; i < n; ++i) fp[(size_t)i] = fingerprint(hk[(size_t)i], g.bits);
  const uint64_t k64 = (uint64_t)lp > g.cap ? (uint64_t)lp - g.cap : 1u;
  const size_t keep0 = k64 >= (uint64_t)n ? 0 : (size_t)n - (size_t)k64;
  std::vector<uint64_t> srt(fp);
  std::nth_element(srt.begin(), srt.begin() + (ptrdiff_t)keep0, srt.end());
  const uint64_t cut = srt[keep0];  // keys with fingerprints >= cut are left out
  std::vector<uint32_t> kept;
  kept.reserve(keep0);
  for (int i = 0; i < n; ++i)
    if (fp[(size_t)i] < cut) kept.push_back(hk[(size_t)i]);
  if (n_failed) *n_failed = (uint32_t)((size_t)n - kept.size());
  uint32_t* dk = nullptr;
  if (!kept.empty()) {
    GQF_HIP(hipMalloc((void**)&dk, kept.size() * 4));
    if (hipMemcpy(dk, kept.data(), kept.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(dk);
      set_error("hipMemcpy of the kept keys failed");
      return GQF_ERR_HIP;
    }
    float t1 = 0.f;
    const gqf_status s2 = rsqf_insert_call(cqf, (int)kept.size(), dk, nullptr, stream, &t1, nullptr);
    (void)hipFree(dk);
    if (s2 != GQF_OK) return s2;
  }
  if (d_out) {
    CallScope cs;
    GQF_TRY(cs.begin(stream));
    launch_rlookup(cs.stream, cqf, (uint32_t)n, d_keys, reinterpret_cast<uint32_t*>(d_out));
    GQF_LAUNCH(k_rk_mark, dim3(((uint32_t)n + 255) / 256), dim3(256), 0, cs.stream, d_keys, (uint32_t)n, g.bits, cut,
               reinterpret_cast<uint32_t*>(d_out));
    GQF_TRY(cs.end(nullptr, false));
  }
  if (ms)
    *ms = t0 + std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return GQF_OK;
}

gqf_status gqf_rsqf_delete(gqf_rsqf cqf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms) {
  GQF_TRY(rsqf_args(cqf, n, d_keys));
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  if (cqf.qbits + cqf.rbits <= 32) GQF_TRY(rsqf_delete_t<uint32_t>(cqf, (uint32_t)n, d_keys, cs));
  else GQF_TRY(rsqf_delete_t<uint64_t>(cqf, (uint32_t)n, d_keys, cs));
  return cs.end(ms, true);
}

gqf_status gqf_rsqf_lookup(gqf_rsqf cqf, int sorted, int n, const uint32_t* d_keys, int32_t* d_out, gqf_stream stream,
                           float* ms) {
  GQF_TRY(rsqf_args(cqf, n, d_keys));
  if (n > 0 && !d_out) {
    set_error("output array is NULL");
    return GQF_ERR_INVALID;
  }
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  uint32_t* out = reinterpret_cast<uint32_t*>(d_out);
  if (n > 0) {
    // sorted or not, the same per-key walk: sorting the keys first costs
    // more than the locality it buys on MI355X (measured up to 1.6 GB tables)
    (void)sorted;
    launch_rlookup(cs.stream, cqf, (uint32_t)n, d_keys, out);
  }
  return cs.end(ms, false);
}

size_t gqf_rsqf_workspace_bytes(gqf_rsqf cqf, int op, int n, int sorted) {
  if (rsqf_check(cqf) != GQF_OK || n < 0) return 0;
  const RsqfGeom g = rgeom(cqf);
  const bool wide = g.bits > 32;
  const uint32_t un = (uint32_t)n;
  (void)sorted;
  if (op == GQF_OP_LOOKUP) return 0;
  if (op == GQF_OP_INSERT) return wide ? rsqf_insert_ws<uint64_t>(g, un) : rsqf_insert_ws<uint32_t>(g, un);
  if (op == GQF_OP_DELETE) return wide ? rsqf_delete_ws<uint64_t>(g, un) : rsqf_delete_ws<uint32_t>(g, un);
  return 0;
}

gqf_status gqf_rsqf_count(gqf_rsqf cqf, int n, const uint32_t* d_keys, uint32_t* d_counts, gqf_stream stream,
                          float* ms) {
  GQF_TRY(rsqf_args(cqf, n, d_keys));
  if (n > 0 && !d_counts) {
    set_error("output array is NULL");
    return GQF_ERR_INVALID;
  }
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  if (n > 0) {
    DeviceState& ds = *cs.ds;
    const uint32_t nb = cqf.numBlocks;
    GQF_TRY(ws_reserve(ds, aligned_bytes((size_t)nb * 8) + 512 + scan_rows_bytes(nb, 2)));
    Arena ar(ds.ws, ds.ws_cap);
    uint32_t* P = ar.take<uint32_t>((size_t)nb * 2);
    uint32_t* tot = ar.take<uint32_t>(2);
    if (ar.used > ar.cap) {
      set_error("internal: count workspace overflow");
      return GQF_ERR_INVALID;
    }
    const RsqfDev d = rdev(cqf);
    GQF_LAUNCH_B((double)nb * 16 + (double)nb * 8, k_rc_pop, dim3((nb + 255) / 256), dim3(256), 0, cs.stream, d.meta,
                 nb, P);
    GQF_TRY(scan_rows<2>(cs.stream, ar, P, nb, tot));
    const dim3 grid(((uint32_t)n + 255) / 256);
    if (cqf.rbits == 8)
      GQF_LAUNCH_B((double)n * (8 + 128), k_rsqf_count<8>, grid, dim3(256), 0, cs.stream, d, cqf.qbits, P, tot, d_keys,
                   (uint32_t)n, d_counts);
    else
      GQF_LAUNCH_B((double)n * (8 + 128), k_rsqf_count<0>, grid, dim3(256), 0, cs.stream, d, cqf.qbits, P, tot, d_keys,
                   (uint32_t)n, d_counts);
  }
  return cs.end(ms, false);
}

// Device layout <-> reference packed blocks, converted on the host.
gqf_status gqf_rsqf_dump(gqf_rsqf cqf, void* host_ref_layout) {
  GQF_TRY(rsqf_check(cqf));
  if (!host_ref_layout) return GQF_ERR_INVALID;
  const uint32_t nb = cqf.numBlocks, r = cqf.rbits;
  std::vector<uint8_t> dev(rsqf_dev_bytes(cqf.qbits, r));
  GQF_HIP(hipMemcpy(dev.data(), cqf.blocks, dev.size(), hipMemcpyDeviceToHost));
  const size_t bb = rsqf_ref_block_bytes(r), mb = rsqf_meta_bytes(nb);
  uint8_t* out = reinterpret_cast<uint8_t*>(host_ref_layout);
  for (uint32_t b = 0; b < nb; ++b) {
    const uint8_t* m = dev.data() + (size_t)b * 32;
    uint8_t* o = out + (size_t)b * bb;
    o[0] = m[16];
    memcpy(o + 1, m, 8);
    memcpy(o + 9, m + 8, 8);
    memcpy(o + 17, dev.data() + mb + (size_t)b * 8 * r, 8 * (size_t)r);
  }
  return GQF_OK;
}

gqf_status gqf_rsqf_load(gqf_rsqf cqf, const void* host_ref_layout) {
  GQF_TRY(rsqf_check(cqf));
  if (!host_ref_layout) return GQF_ERR_INVALID;
  const uint32_t nb = cqf.numBlocks, r = cqf.rbits;
  std::vector<uint8_t> dev(rsqf_dev_bytes(cqf.qbits, r), 0);
  const size_t bb = rsqf_ref_block_bytes(r), mb = rsqf_meta_bytes(nb);
  const uint8_t* in = reinterpret_cast<const uint8_t*>(host_ref_layout);
  for (uint32_t b = 0; b < nb; ++b) {
    uint8_t* m = dev.data() + (size_t)b * 32;
    const uint8_t* i = in + (size_t)b * bb;
    memcpy(m, i + 1, 8);
    memcpy(m + 8, i + 9, 8);
    m[16] = i[0];
    memcpy(dev.data() + mb + (size_t)b * 8 * r, i + 17, 8 * (size_t)r);
  }
  GQF_HIP(hipMemcpy(cqf.blocks, dev.data(), dev.size(), hipMemcpyHostToDevice));
  return GQF_OK;
}

}  // extern "C"
