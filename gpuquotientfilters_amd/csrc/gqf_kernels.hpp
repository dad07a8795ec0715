// gqf_kernels.hpp -- device primitives shared by the SQF and RSQF engines.
//
//   * fingerprint(): Normal_APHash (quotientFilter.cu:134-149 / RSQF.cu:174-189)
//     with the wide extension for q + r > 32;
//   * an LDS-staged LSD radix sort, chunked reduce-then-scan (per pass: chunk
//     digit histograms, a row scan of them, then a downsweep that ranks each
//     4096-key sub-tile with wave64 ballots and writes runs of equal digits;
//     no inter-workgroup waits), optionally hashing raw keys on the fly and
//     carrying a uint32 index -- replaces thrust::sort / sort_by_key
//     (quotientFilter.cu:295, 304, 770, 963, 1475; RSQF.cu:433, 442, 913);
//   * the row scans the decoders use for per-region counts;
//   * merge-path splits with a source tag -- replaces mgpu::merge
//     (quotientFilter.cu:1651, 1800);
//   * delete marking (multiset difference, one stored copy per delete key).
// The canonical-layout scan (loc_i = max(fq_i, loc_{i-1} + 1), a
// reduce-then-scan over chunk states fused with the keep rules) lives in
// gqf_rebuild.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include "gqf_runtime.hpp"

namespace gqf {

// ------------------------------------------------------------------ helpers
__host__ __device__ __forceinline__ uint64_t mask64(unsigned n) {
  return n >= 64 ? ~0ull : ((1ull << n) - 1ull);
}

// Normal_APHash rounds over the big-endian bytes of key, from state h.
__host__ __device__ __forceinline__ uint32_t ap_rounds(uint32_t key, uint32_t h) {
  h ^= (h << 7) ^ (key >> 24) ^ (h >> 3);
  h ^= ~((h << 11) ^ ((key >> 16) & 0xFFu) ^ (h >> 5));
  h ^= (h << 7) ^ ((key >> 8) & 0xFFu) ^ (h >> 3);
  h ^= ~((h << 11) ^ (key & 0xFFu) ^ (h >> 5));
  return h;
}

// MurmurHash3 64-bit finalizer (fmix64; Appleby, public domain).
__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Fingerprint of width bits = q + r.  bits <= 31: AP(key) mod 2^bits (the
// reference); bits == 32: AP(key); bits > 32 (undefined in the reference):
// fmix64(AP(key) << 32 | key) truncated to bits -- uniform, because the input
// word is injective in the key.
__host__ __device__ __forceinline__ uint64_t fingerprint(uint32_t key, unsigned bits) {
  uint32_t lo = ap_rounds(key, 0xAAAAAAAAu);
  if (bits <= 32) return bits == 32 ? (uint64_t)lo : (uint64_t)(lo & ((1u << bits) - 1u));
  return fmix64(((uint64_t)lo << 32) | key) & mask64(bits);
}

// Position of the k-th set bit of x (k >= 1, 0-based position), found by
// halving with popcounts; 64 if x has fewer than k set bits.
__host__ __device__ __forceinline__ uint32_t select_bit64(uint64_t x, uint32_t k) {
  if (k == 0 || k > (uint32_t)__builtin_popcountll(x)) return 64u;
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t half = 32; half >= 1; half >>= 1) {
    const uint64_t lowm = half == 64 ? ~0ull : ((1ull << half) - 1ull);
    const uint32_t c = (uint32_t)__builtin_popcountll(x & lowm);
    if (k > c) {
      k -= c;
      x >>= half;
      pos += half;
    }
  }
  return pos;
}

// ---------------------------------------------------------------- wave64 scans
// Inclusive scans across the 64 lanes of a wave with DPP moves: row_shr 1, 2,
// 4, 8 inside each 16-lane row, then row_bcast:15 (rows 1, 3) and
// row_bcast:31 (rows 2, 3); wave_shr:1 turns an inclusive scan into an
// exclusive one.  A lane with no source lane receives the identity.  Every
// lane of the wave must be active.
constexpr int DPP_ROW_SHR = 0x110;  // | shift (1..15)
constexpr int DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t ident, uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)x, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t ident, uint64_t x) {
  const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)ident, (uint32_t)x);
  const uint32_t hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(ident >> 32), (uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_incl_add(uint32_t x) {
  x += dpp_u32<DPP_ROW_SHR | 1>(0u, x);
  x += dpp_u32<DPP_ROW_SHR | 2>(0u, x);
  x += dpp_u32<DPP_ROW_SHR | 4>(0u, x);
  x += dpp_u32<DPP_ROW_SHR | 8>(0u, x);
  x += dpp_u32<DPP_ROW_BCAST15, 0xA>(0u, x);
  x += dpp_u32<DPP_ROW_BCAST31, 0xC>(0u, x);
  return x;
}
__device__ __forceinline__ uint64_t wave_incl_add64(uint64_t x) {
  x += dpp_u64<DPP_ROW_SHR | 1>(0ull, x);
  x += dpp_u64<DPP_ROW_SHR | 2>(0ull, x);
  x += dpp_u64<DPP_ROW_SHR | 4>(0ull, x);
  x += dpp_u64<DPP_ROW_SHR | 8>(0ull, x);
  x += dpp_u64<DPP_ROW_BCAST15, 0xA>(0ull, x);
  x += dpp_u64<DPP_ROW_BCAST31, 0xC>(0ull, x);
  return x;
}
__device__ __forceinline__ uint32_t wave_read(uint32_t x, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

// Exclusive scan of one uint32 per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block256_excl_scan(uint32_t x, uint32_t* s_tmp4, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_add(x);
  if (lane == 63) s_tmp4[w] = inc;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t v = s_tmp4[i];
    pre += (i < w) ? v : 0u;
    tot += v;
  }
  __syncthreads();
  if (total) *total = tot;
  return pre + inc - x;
}

// Exclusive scan of one uint32 per thread over a block of T threads (T a
// multiple of 64); s_tmp: T/64 words.  Contains two barriers.
template <int T>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_tmp) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_add(x);
  if (lane == 63) s_tmp[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
#pragma unroll
  for (int i = 0; i < T / 64; ++i) pre += (i < w) ? s_tmp[i] : 0u;
  __syncthreads();
  return pre + inc - x;
}

// Same for uint64 (used with packed counters).
__device__ __forceinline__ uint64_t block256_excl_scan64(uint64_t x, uint64_t* s_tmp4, uint64_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t inc = wave_incl_add64(x);
  if (lane == 63) s_tmp4[w] = inc;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint64_t v = s_tmp4[i];
    pre += (i < w) ? v : 0ull;
    tot += v;
  }
  __syncthreads();
  if (total) *total = tot;
  return pre + inc - x;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

[[maybe_unused]] static __global__ void k_reset_counters(uint32_t* ctr, int from, int to) {
  int i = from + threadIdx.x;
  if (i < to) ctr[i] = 0;
}

template <typename T>
inline void launch_fill(hipStream_t s, T* p, uint64_t n, T v) {
  if (!n) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  GQF_LAUNCH(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

// ====================================================================== radix sort
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 4096 keys per tile
constexpr int SORT_MAX_PASSES = 8;

// Lanes of the wave holding the same 8-bit digit (8 ballots, all 64 lanes active).
__device__ __forceinline__ uint64_t match_digit(uint32_t d) {
  uint64_t m = ~0ull;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    uint64_t bb = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bb : ~bb;
  }
  return m;
}

// FNVhashGPU (quotientFilter.cu:117-132): FNV-1 over the key's big-endian
// bytes (multiply by the FNV prime, then xor the byte), from the FNV offset
// basis.  The SQF's alternative hash (GQF_HASH_FNV).
__host__ __device__ __forceinline__ uint32_t fnv_rounds(uint32_t key) {
  uint32_t h = 2166136261u;
  h = (h * 16777619u) ^ (key >> 24);
  h = (h * 16777619u) ^ ((key >> 16) & 0xFFu);
  h = (h * 16777619u) ^ ((key >> 8) & 0xFFu);
  h = (h * 16777619u) ^ (key & 0xFFu);
  return h;
}

// Flag in a fingerprint width argument: hash with FNV instead of AP.
constexpr uint32_t FP_FNV = 0x100u;

// Fingerprint when the result fits 32 bits (q + r <= 32), bits = q + r
// (| FP_FNV): the hash mod 2^bits, as the reference's `hash % tableSize`.
__host__ __device__ __forceinline__ uint32_t fingerprint32(uint32_t key, unsigned bits) {
  const unsigned b = bits & 0xFFu;
  const uint32_t h = (bits & FP_FNV) ? fnv_rounds(key) : ap_rounds(key, 0xAAAAAAAAu);
  return b >= 32 ? h : (h & ((1u << b) - 1u));
}

template <typename K, bool HASH>
__device__ __forceinline__ K sort_load_key(const void* in, uint64_t idx, uint32_t fp_bits) {
  if (HASH) {
    const uint32_t key = reinterpret_cast<const uint32_t*>(in)[idx];
    // a 32-bit key type means q + r <= 32: skip the wide path entirely
    if (sizeof(K) == 4) return (K)fingerprint32(key, fp_bits);
    return (K)fingerprint(key, fp_bits);
  }
  return reinterpret_cast<const K*>(in)[idx];
}

// Chunked reduce-then-scan LSD radix sort.  The input is split into G
// contiguous chunks (one workgroup each, ~2K chunks so every CU stays busy);
// per 8-bit digit pass:
//   upsweep   -- each chunk's digit histogram (one read of the keys),
//   row scan  -- exclusive scan of every digit's chunk counts and the digit
//                totals,
//   downsweep -- the digit bases (a scan of the 256 totals, in every
//                workgroup), then each chunk walks its keys in 4096-key
//                sub-tiles in order,
//                ranks them stably with wave64 ballots, stages them in LDS by
//                digit and writes runs of equal digits (second read + write).
// No workgroup waits on another and no counter atomics are used: on MI355X a
// same-address device atomic serializes at ~11 ns and a cross-workgroup
// publish/observe round trip costs ~0.4 us, which made a single-pass
// (look-back) sort latency-bound.
struct SortPlan {
  uint32_t G;      // chunks
  uint32_t chunk;  // keys per chunk (multiple of SORT_TILE)
};
inline SortPlan sort_plan(uint32_t n) {
  SortPlan p;
  const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
  if (tiles == 0) {  // empty input: one (idle) chunk
    p.chunk = SORT_TILE;
    p.G = 1;
    return p;
  }
  const uint32_t G0 = tiles < 2048 ? tiles : 2048u;
  // an even number of 4096-key tiles: whole 8192-key downsweep sub-tiles
  const uint32_t per = ((tiles + G0 - 1) / G0 + 1) & ~1u;
  p.chunk = per * SORT_TILE;
  p.G = (uint32_t)(((uint64_t)n + p.chunk - 1) / p.chunk);
  return p;
}

template <typename K, bool HASH>
__global__ __launch_bounds__(256) void k_rs_upsweep(const void* __restrict__ in, uint32_t n, uint32_t fp_bits,
                                                    uint32_t shift, uint32_t chunk, uint32_t G,
                                                    uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4 * 256];
  const int tid = threadIdx.x, wid = tid >> 6;
  for (int i = tid; i < 4 * 256; i += 256) h[i] = 0;
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * chunk;
  const uint64_t hi = lo + chunk < n ? lo + chunk : (uint64_t)n;
  for (uint64_t t0 = lo; t0 < hi; t0 += SORT_TILE) {
    // all SORT_ITEMS loads are issued before any is used
    K k[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint64_t idx = t0 + (uint64_t)i * 256 + tid;
      k[i] = idx < hi ? sort_load_key<K, HASH>(in, idx, fp_bits) : (K)0;
    }
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint64_t idx = t0 + (uint64_t)i * 256 + tid;
      if (idx < hi) atomicAdd(&h[wid * 256 + ((uint32_t)(k[i] >> shift) & 255u)], 1u);
    }
  }
  __syncthreads();
  hist[(uint64_t)tid * G + blockIdx.x] = h[tid] + h[256 + tid] + h[512 + tid] + h[768 + tid];
}

// Row d (one workgroup per digit): exclusive scan of the G chunk counts in
// place; the row total goes to totals[d].  crowd (may be null): a word of
// host-coherent memory set to 1 when a row total exceeds `limit`.  clear
// (may be null): a word set to zero.
[[maybe_unused]] static __global__ __launch_bounds__(1024) void k_rs_rowscan(uint32_t* __restrict__ hist, uint32_t G,
                                                            uint32_t* __restrict__ totals,
                                                            uint32_t* __restrict__ crowd, uint32_t limit,
                                                            uint32_t* __restrict__ clear) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (clear && blockIdx.x == 0 && tid == 0) *clear = 0u;
  uint32_t* row = hist + (uint64_t)blockIdx.x * G;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < G; b += 1024) {
    const uint32_t i = b + tid;
    const uint32_t x = i < G ? row[i] : 0u;
    const uint32_t inc = wave_incl_add(x);
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      pre += w < wid ? s_w[w] : 0u;
      tot += s_w[w];
    }
    const uint32_t carry = s_carry;
    if (i < G) row[i] = carry + pre + inc - x;
    __syncthreads();
    if (tid == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (tid == 0) {
    totals[blockIdx.x] = s_carry;
    if (crowd && s_carry > limit) __hip_atomic_store(crowd, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------- LDS tile sort
// A sub-tile of SORT_TILE entries is sorted stably by an 8-bit digit in LDS
// with two 4-bit counting passes (low nibble, then high nibble), in blocked
// order: thread t holds tile positions 16t .. 16t+15.  Per pass:
//   1. each thread counts its 16 nibbles in one 64-bit register of 4-bit
//      fields; the field's value before an entry is the entry's rank among
//      the thread's entries of that nibble (a field reaches 16 only when all
//      16 entries share the nibble, which the 16th entry's rank of 15 shows);
//   2. the 16 x 256 counts, nibble-major and thread-minor, are scanned;
//   3. every entry goes to prefix(nibble, t) + rank.
// Positions of equal keys never swap, so padding entries (all-ones keys at
// the end of a short tile) keep the last positions of the last digit.  This
// replaces per-entry matching of the 8 digit bits with wave ballots, whose
// ~90 VALU instructions per entry made the downsweep ALU-bound.
//
// In LDS the 16-byte chunks of every 16 consecutive entries are XOR-permuted
// (ts_swz) so that threads reading their own 16 entries -- 64 or 128 bytes
// apart -- hit different banks; rows read by consecutive lanes stay whole.
template <int E>
__device__ __forceinline__ uint32_t ts_swz(uint32_t p) {
  static_assert(E == 4 || E == 8, "4- or 8-byte entries");
  if (E == 4) return p ^ (((p >> 6) & 3u) << 2);
  return p ^ (((p >> 5) & 7u) << 1);
}

// 16 bytes from a 4-byte aligned address: one global_load_dwordx4 (gfx950
// multi-dword loads need dword alignment only), so chunks that start at any
// key still load whole groups.
__device__ __forceinline__ uint4 ld16_dw(const uint32_t* p) {
  uint4 x;
  __builtin_memcpy(&x, p, 16);
  return x;
}

// Thread t's 16 consecutive entries [16t, 16t+16) of the sub-tile at t0
// (tile_n of them valid; the rest become all-ones padding); whole groups
// load as 16-byte words.
template <typename K, bool HASH>
__device__ __forceinline__ void ts_load_keys(const void* __restrict__ in, uint64_t t0, uint32_t tile_n,
                                             uint32_t fp_bits, K (&k)[SORT_ITEMS]) {
  const uint32_t i0 = threadIdx.x * SORT_ITEMS;
  if (HASH || sizeof(K) == 4) {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(in) + t0 + i0;
    uint32_t w[SORT_ITEMS];
    if (i0 + SORT_ITEMS <= tile_n) {
#pragma unroll
      for (int c = 0; c < SORT_ITEMS / 4; ++c) {
        const uint4 x = ld16_dw(p + 4 * c);
        w[4 * c] = x.x;
        w[4 * c + 1] = x.y;
        w[4 * c + 2] = x.z;
        w[4 * c + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < SORT_ITEMS; ++i) w[i] = i0 + i < tile_n ? p[i] : 0u;
    }
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      K x;
      if (HASH) x = sizeof(K) == 4 ? (K)fingerprint32(w[i], fp_bits) : (K)fingerprint(w[i], fp_bits);
      else x = (K)w[i];
      k[i] = i0 + i < tile_n ? x : (K)~(K)0;
    }
  } else {
    const K* p = reinterpret_cast<const K*>(in) + t0 + i0;
    if (i0 + SORT_ITEMS <= tile_n) {
#pragma unroll
      for (int c = 0; c < SORT_ITEMS / 2; ++c) {
        const uint4 x = ld16_dw(reinterpret_cast<const uint32_t*>(p + 2 * c));
        k[2 * c] = (K)(((uint64_t)x.y << 32) | x.x);
        k[2 * c + 1] = (K)(((uint64_t)x.w << 32) | x.z);
      }
    } else {
#pragma unroll
      for (int i = 0; i < SORT_ITEMS; ++i) k[i] = i0 + i < tile_n ? p[i] : (K)~(K)0;
    }
  }
}

// Values of the same positions; without an input array, the input index.
__device__ __forceinline__ void ts_load_vals(const uint32_t* __restrict__ in_vals, uint64_t t0, uint32_t tile_n,
                                             uint32_t (&v)[SORT_ITEMS]) {
  const uint32_t i0 = threadIdx.x * SORT_ITEMS;
  if (!in_vals) {
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) v[i] = (uint32_t)(t0 + i0 + i);
    return;
  }
  const uint32_t* p = in_vals + t0 + i0;
  if (i0 + SORT_ITEMS <= tile_n) {
#pragma unroll
    for (int c = 0; c < SORT_ITEMS / 4; ++c) {
      const uint4 x = ld16_dw(p + 4 * c);
      v[4 * c] = x.x;
      v[4 * c + 1] = x.y;
      v[4 * c + 2] = x.z;
      v[4 * c + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) v[i] = i0 + i < tile_n ? p[i] : 0u;
  }
}

// Step 1 of a counting pass over the nibble at bit sh: the entries' ranks
// within the thread (4 bits each, entry i at bits 4i of *rk) and the thread's
// 16 counts, written to s_cnt[j*T + t] (T threads in the block).
template <int T, typename K>
__device__ __forceinline__ void ts_count(const K (&k)[SORT_ITEMS], uint32_t sh, uint64_t& rk, uint16_t* s_cnt) {
  uint64_t h = 0;
  uint32_t last = 0, r = 0;
  rk = 0;
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint32_t d = (uint32_t)(k[i] >> sh) & 15u;
    r = (uint32_t)(h >> (4 * d)) & 15u;
    rk |= (uint64_t)r << (4 * i);
    h += 1ull << (4 * d);
    last = d;
  }
  const bool full = r == 15u;  // the 16th entry saw 15 before it: all 16 share nibble `last`
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint32_t c = full ? (j == (int)last ? 16u : 0u) : (uint32_t)(h >> (4 * j)) & 15u;
    s_cnt[j * T + t] = (uint16_t)c;
  }
}

// Step 2: thread t scans the 16 counts s_cnt[16t .. 16t+15] (extra, a value
// scanned alongside in the high half of a 32-bit word: both halves stay below
// 2^16) and writes exclusive prefixes back.  Returns the exclusive prefix of
// `extra` over the block.  Contains two barriers.
template <int T>
__device__ __forceinline__ uint32_t ts_scan(uint16_t* s_cnt, uint32_t extra, uint32_t* s_tmp) {
  const uint32_t t = threadIdx.x;
  uint4* row = reinterpret_cast<uint4*>(s_cnt + 16 * t);
  uint4 a = row[0], b = row[1];
  uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) sum += (w[j] & 0xFFFFu) + (w[j] >> 16);
  const uint32_t ex = block_excl_scan<T>(sum | (extra << 16), s_tmp);
  uint32_t run = ex & 0xFFFFu;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t lo = run;
    run += w[j] & 0xFFFFu;
    const uint32_t hi = run;
    run += w[j] >> 16;
    w[j] = lo | (hi << 16);
  }
  row[0] = make_uint4(w[0], w[1], w[2], w[3]);
  row[1] = make_uint4(w[4], w[5], w[6], w[7]);
  return ex >> 16;
}

// Step 3: every entry to its position (keys and values, swizzled).
template <int T, typename K, bool VALS>
__device__ __forceinline__ void ts_scatter(const K (&k)[SORT_ITEMS], const uint32_t (&v)[SORT_ITEMS], uint32_t sh,
                                           uint64_t rk, const uint16_t* s_cnt, K* s_keys, uint32_t* s_vals) {
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint32_t d = (uint32_t)(k[i] >> sh) & 15u;
    const uint32_t pos = (uint32_t)s_cnt[d * T + t] + ((uint32_t)(rk >> (4 * i)) & 15u);
    s_keys[ts_swz<sizeof(K)>(pos)] = k[i];
    if (VALS) s_vals[ts_swz<4>(pos)] = v[i];
  }
}

// Thread t's 16 sorted entries [16t, 16t+16) back from LDS.
template <typename K, bool VALS>
__device__ __forceinline__ void ts_read_blocked(const K* s_keys, const uint32_t* s_vals, K (&k)[SORT_ITEMS],
                                                uint32_t (&v)[SORT_ITEMS]) {
  const uint32_t b = threadIdx.x * SORT_ITEMS;
  if (sizeof(K) == 4) {
#pragma unroll
    for (int c = 0; c < SORT_ITEMS / 4; ++c) {
      const uint4 x = *reinterpret_cast<const uint4*>(s_keys + ts_swz<4>(b + 4 * c));
      k[4 * c] = (K)x.x;
      k[4 * c + 1] = (K)x.y;
      k[4 * c + 2] = (K)x.z;
      k[4 * c + 3] = (K)x.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < SORT_ITEMS / 2; ++c) {
      const uint4 x = *reinterpret_cast<const uint4*>(s_keys + ts_swz<8>(b + 2 * c));
      k[2 * c] = (K)(((uint64_t)x.y << 32) | x.x);
      k[2 * c + 1] = (K)(((uint64_t)x.w << 32) | x.z);
    }
  }
  if (VALS) {
#pragma unroll
    for (int c = 0; c < SORT_ITEMS / 4; ++c) {
      const uint4 x = *reinterpret_cast<const uint4*>(s_vals + ts_swz<4>(b + 4 * c));
      v[4 * c] = x.x;
      v[4 * c + 1] = x.y;
      v[4 * c + 2] = x.z;
      v[4 * c + 3] = x.w;
    }
  }
}

// Downsweep of one 8-bit pass: each chunk walks its sub-tiles in order,
// sorts each by the digit in LDS (the tile sort above) and writes every
// digit's run to the digit base + the chunk's offset + the runs of earlier
// sub-tiles.  The sub-tile's digit counts come from LDS atomics on per-wave
// histograms during the first counting pass.
//   Sub-tiles hold 16 keys per thread, T threads: 32-bit keys without
// values take 512 threads (8192 keys), so a digit's run per sub-tile is 32
// keys, one 128-byte line -- runs of 16 keys cap the pattern at 3.7 TB/s,
// runs of 32 at 4.8 (a probe of the bare pattern); 64-bit keys reach whole
// lines with 256 threads.
//   Occupancy: 32-bit keys, 512 threads: 58 KB of LDS and 128 VGPRs, two
// workgroups (16 waves) per CU; 64-bit keys or values, 256 threads: 46 KB,
// three; both: 62 KB, two.
template <typename K, bool VALS>
constexpr int ds_threads() { return (sizeof(K) == 4 && !VALS) ? 512 : 256; }

// The downsweep of one chunk [c0, c1): sub-tile by sub-tile, sorted by the
// digit in LDS and written to s_run[digit] + the runs of earlier sub-tiles
// (s_run: the chunk's output base of every digit; s_hist: all zeros).
template <typename K, bool VALS, bool HASH, int T>
__device__ __forceinline__ void ds_chunk(const void* __restrict__ in_keys, const uint32_t* __restrict__ in_vals,
                                         K* __restrict__ out_keys, uint32_t* __restrict__ out_vals, uint32_t fp_bits,
                                         uint32_t shift, uint64_t c0, uint64_t c1, K* s_keys, uint32_t* s_vals,
                                         uint16_t* s_cnt, uint32_t* s_hist, uint32_t* s_run, uint32_t* s_base,
                                         uint32_t* s_tmp) {
  constexpr int TILE = T * SORT_ITEMS;
  constexpr int W = T / 64;
  const uint32_t tid = threadIdx.x, wid = tid >> 6;
  const bool dig = tid < 256;
  K k[SORT_ITEMS];
  uint32_t v[SORT_ITEMS];
  uint64_t rk;
  if (c0 < c1) {
    const uint32_t tn = (uint32_t)(c1 - c0 < (uint64_t)TILE ? c1 - c0 : (uint64_t)TILE);
    ts_load_keys<K, HASH>(in_keys, c0, tn, fp_bits, k);
    if (VALS) ts_load_vals(in_vals, c0, tn, v);
  }
  for (uint64_t t0 = c0; t0 < c1; t0 += TILE) {
    const uint32_t tile_n = (uint32_t)(c1 - t0 < (uint64_t)TILE ? c1 - t0 : (uint64_t)TILE);
    // pass 1 (low nibble) and the sub-tile's digit histogram
    ts_count<T>(k, shift, rk, s_cnt);
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) atomicAdd(&s_hist[wid * 256 + ((uint32_t)(k[i] >> shift) & 255u)], 1u);
    __syncthreads();
    uint32_t total = 0;
    if (dig) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        total += s_hist[w * 256 + tid];
        s_hist[w * 256 + tid] = 0;  // the next sub-tile's counts start after several barriers
      }
    }
    const uint32_t lstart = ts_scan<T>(s_cnt, total, s_tmp);
    if (dig) s_base[tid] = s_run[tid] - lstart;  // output of sorted position p with digit tid: s_base + p
    __syncthreads();
    ts_scatter<T, K, VALS>(k, v, shift, rk, s_cnt, s_keys, s_vals);
    __syncthreads();
    ts_read_blocked<K, VALS>(s_keys, s_vals, k, v);
    // pass 2 (high nibble)
    ts_count<T>(k, shift + 4, rk, s_cnt);
    __syncthreads();
    (void)ts_scan<T>(s_cnt, 0u, s_tmp);
    __syncthreads();
    ts_scatter<T, K, VALS>(k, v, shift + 4, rk, s_cnt, s_keys, s_vals);
    __syncthreads();
    // the next sub-tile's loads go out before this one's stores
    const uint64_t tn0 = t0 + TILE;
    if (tn0 < c1) {
      const uint32_t tn = (uint32_t)(c1 - tn0 < (uint64_t)TILE ? c1 - tn0 : (uint64_t)TILE);
      ts_load_keys<K, HASH>(in_keys, tn0, tn, fp_bits, k);
      if (VALS) ts_load_vals(in_vals, tn0, tn, v);
    }
    // runs of equal digits to their outputs; consecutive lanes, consecutive positions
#pragma unroll 4
    for (int j = 0; j < SORT_ITEMS; ++j) {
      const uint32_t p = tid + (uint32_t)j * T;
      if (p < tile_n) {
        const K key = s_keys[ts_swz<sizeof(K)>(p)];
        const uint32_t gpos = s_base[(uint32_t)(key >> shift) & 255u] + p;
        out_keys[gpos] = key;
        if (VALS) out_vals[gpos] = s_vals[ts_swz<4>(p)];
      }
    }
    if (dig) s_run[tid] += total;
    // no barrier: the next sub-tile writes s_base and s_keys only after its
    // first barrier, which every thread reaches after these reads
  }
}

// (the second bound is waves per SIMD: workgroups per CU times T/256)
template <typename K, bool VALS, bool HASH>
__global__ __launch_bounds__((ds_threads<K, VALS>()), (sizeof(K) == 4 ? (VALS ? 3 : 4) : (VALS ? 2 : 3))) void
k_rs_downsweep(const void* __restrict__ in_keys, const uint32_t* __restrict__ in_vals, K* __restrict__ out_keys,
               uint32_t* __restrict__ out_vals, uint32_t n, uint32_t fp_bits, uint32_t shift, uint32_t chunk,
               uint32_t G, const uint32_t* __restrict__ hist, const uint32_t* __restrict__ digit_totals) {
  constexpr int T = ds_threads<K, VALS>();
  constexpr int W = T / 64;
  __shared__ __align__(16) K s_keys[T * SORT_ITEMS];
  __shared__ __align__(16) uint32_t s_vals[VALS ? T * SORT_ITEMS : 4];
  __shared__ __align__(16) uint16_t s_cnt[16 * T];
  __shared__ uint32_t s_hist[W * 256];
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_tmp[W];
  const uint32_t tid = threadIdx.x;
  const bool dig = tid < 256;  // the threads that own a digit
#pragma unroll
  for (int w = 0; w < W * 256 / T; ++w) s_hist[w * T + tid] = 0;
  // digit bases: exclusive scan of the 256 digit totals (every workgroup
  // does it; it saves a launch), plus this chunk's offset
  {
    const uint32_t base = block_excl_scan<T>(dig ? digit_totals[tid] : 0u, s_tmp);
    if (dig) s_run[tid] = base + hist[(uint64_t)tid * G + blockIdx.x];
  }
  const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t c1 = c0 + chunk < n ? c0 + chunk : (uint64_t)n;
  ds_chunk<K, VALS, HASH, T>(in_keys, in_vals, out_keys, out_vals, fp_bits, shift, c0, c1, s_keys, s_vals, s_cnt,
                             s_hist, s_run, s_base, s_tmp);
}

// Every pass of a sort of at most SORT_TILE keys in one workgroup: the keys
// (and values) stay in registers and LDS between passes; each pass ranks them
// stably as k_rs_downsweep does (ballots within a wave, then waves in order)
// and scatters them within LDS.  Key j sits at position wid*1024 + i*64 +
// lane of thread (wid, lane), item i, before and after every pass; padding
// keys (all ones) take the last digit and rank after every real key.
template <typename K, bool VALS, bool HASH>
__global__ __launch_bounds__(SORT_THREADS) void k_rs_small(const void* __restrict__ in_keys,
                                                           const uint32_t* __restrict__ in_vals,
                                                           K* __restrict__ out_keys, uint32_t* __restrict__ out_vals,
                                                           uint32_t n, uint32_t fp_bits, uint32_t lo_bit,
                                                           uint32_t npass) {
  __shared__ uint32_t s_wcnt[4 * 256];
  __shared__ uint32_t s_lstart[256];
  __shared__ uint32_t s_tmp[4];
  __shared__ K s_keys[SORT_TILE];
  __shared__ uint32_t s_vals[VALS ? SORT_TILE : 1];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  K k[SORT_ITEMS];
  uint32_t v[VALS ? SORT_ITEMS : 1];
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint32_t idx = (uint32_t)wid * (SORT_ITEMS * 64) + (uint32_t)i * 64 + lane;
    const bool valid = idx < n;
    k[i] = valid ? sort_load_key<K, HASH>(in_keys, idx, fp_bits) : (K)~(K)0;
    if (VALS) v[i] = in_vals ? (valid ? in_vals[idx] : 0u) : idx;
  }
  for (uint32_t p = 0; p < npass; ++p) {
    const uint32_t shift = lo_bit + 8 * p;
    for (int i = tid; i < 4 * 256; i += SORT_THREADS) s_wcnt[i] = 0;
    __syncthreads();
    uint32_t rk[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
      const uint64_t m = match_digit(d);
      const uint32_t pre = s_wcnt[wid * 256 + d];
      const uint32_t cnt = (uint32_t)__popcll(m);
      const uint32_t below = (uint32_t)__popcll(m & lt);
      if (below == cnt - 1) s_wcnt[wid * 256 + d] = pre + cnt;
      rk[i] = pre + below;
    }
    __syncthreads();
    {
      const uint32_t d = tid;
      const uint32_t w0 = s_wcnt[d], w1 = s_wcnt[256 + d], w2 = s_wcnt[512 + d], w3 = s_wcnt[768 + d];
      s_wcnt[d] = 0;
      s_wcnt[256 + d] = w0;
      s_wcnt[512 + d] = w0 + w1;
      s_wcnt[768 + d] = w0 + w1 + w2;
      s_lstart[d] = block256_excl_scan(w0 + w1 + w2 + w3, s_tmp, nullptr);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
      const uint32_t pos = s_lstart[d] + s_wcnt[wid * 256 + d] + rk[i];
      s_keys[pos] = k[i];
      if (VALS) s_vals[pos] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t pos = (uint32_t)wid * (SORT_ITEMS * 64) + (uint32_t)i * 64 + lane;
      k[i] = s_keys[pos];
      if (VALS) v[i] = s_vals[pos];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint32_t idx = (uint32_t)wid * (SORT_ITEMS * 64) + (uint32_t)i * 64 + lane;
    if (idx < n) {
      out_keys[idx] = k[i];
      if (VALS) out_vals[idx] = v[i];
    }
  }
}

// ---------------------------------------------------------------- segment sort
// A keys-only sort with the stream's context replaces the 8-bit passes of
// the chunked sort by a partition on the top bits and one LDS sort per group
// of keys:
//   * 4097 .. 256 * SEG_GROUP_AVG keys: one chunked pass on the top 8 bits
//     of the key range (a stable partition into 256 segments), then
//     k_rs_seg, one workgroup per segment, sorts the bits below in LDS with
//     every 4-bit counting pass of the tile sort back to back.  4 launches
//     instead of 12 (31-bit fingerprints) or 15 (34-bit ones).
//   * more keys: a second chunked pass over the segments on the next b2 bits
//     (k_rs_upsweep_seg, k_rs_segbase, k_rs_downsweep_seg: each chunk lies in
//     one segment, and the digit bases count within the segment), so that
//     256 << b2 groups of at most ~7600 keys on average remain for k_rs_seg.
//     Two chunked passes and one LDS pass instead of 4 (or 5) chunked passes.
// Within a group every key shares the bits above the sorted range, so the
// LDS sort holds the low 32 bits only (also for 64-bit keys).
//   Skewed keys can crowd a group (the AP hash puts most small integer keys
// under a few top-bit values): the row scan of the last partition raises a
// host-visible flag when a group exceeds SEG_TILE, the host reads it while
// the partition's downsweep runs, and then sorts the partitioned keys with
// the chunked passes instead.
constexpr int SEG_THREADS = 512;
constexpr int SEG_TILE = SEG_THREADS * SORT_ITEMS;  // 8192 keys
constexpr uint32_t SEG_BB = 13;                     // bucket bits of the LDS group sort
constexpr uint32_t SEG_NBK = 1u << SEG_BB;          // its buckets (16-bit counters)
constexpr uint32_t SEG_RANK_MAX = 32;               // largest bucket ranked key by key
constexpr uint32_t SEG_SLOW_GRID = 256;             // workgroups walking the list of groups k_rs_seg left
// Average group size the partitions aim below: the largest of 256 << b2
// groups of uniform keys stays under SEG_TILE (7600 + 5 sigma < 8192).
constexpr uint32_t SEG_GROUP_AVG = 7600;

// b2 of a sort of n keys: the fewest bits for 256 << b2 groups of at most
// SEG_GROUP_AVG keys on average (0: the top-digit segments are the groups;
// 9: too many keys for one more 8-bit pass).
inline uint32_t seg_bits2(uint64_t n) {
  uint32_t b = 0;
  while (b <= 8 && n > ((uint64_t)256u << b) * SEG_GROUP_AVG) ++b;
  return b;
}

// Chunk x of a pass over the 256 segments of a partitioned array (sizes
// totals[0..255], in order), each cut into chunks of `chunk` keys: its range
// [*c0, *c1) and its segment (256 and an empty range past the last chunk).
// s_a, s_b: 257 words of LDS each (segment starts, chunk prefixes).
template <int T>
__device__ __forceinline__ uint32_t seg_chunk(const uint32_t* __restrict__ totals, uint32_t chunk, uint32_t x,
                                              uint64_t* c0, uint64_t* c1, uint32_t* s_a, uint32_t* s_b,
                                              uint32_t* s_tmp) {
  const uint32_t t = threadIdx.x;
  const uint32_t m = t < 256 ? totals[t] : 0u;
  const uint32_t nc = (m + chunk - 1) / chunk;
  const uint32_t st = block_excl_scan<T>(m, s_tmp);
  const uint32_t cp = block_excl_scan<T>(nc, s_tmp);
  if (t < 256) {
    s_a[t] = st;
    s_b[t] = cp;
    if (t == 255) {
      s_a[256] = st + m;
      s_b[256] = cp + nc;
    }
  }
  __syncthreads();
  if (x >= s_b[256]) {
    *c0 = *c1 = 0;
    return 256;
  }
  uint32_t lo = 0, hi = 256;  // the last segment whose first chunk is at or before x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_b[mid] <= x) lo = mid;
    else hi = mid;
  }
  // uniform over the workgroup: kept in scalar registers
  lo = __builtin_amdgcn_readfirstlane(lo);
  const uint32_t a0 = __builtin_amdgcn_readfirstlane(s_a[lo]), a1 = __builtin_amdgcn_readfirstlane(s_a[lo + 1]);
  const uint32_t b0 = __builtin_amdgcn_readfirstlane(s_b[lo]);
  *c0 = a0 + (uint64_t)(x - b0) * chunk;
  *c1 = *c0 + chunk < a1 ? *c0 + chunk : (uint64_t)a1;
  return lo;
}

// Digit histogram of chunk x (blockIdx.x) of the segmented pass, to
// hist2[x * 256 + digit] (zeros past the last chunk).
// cb: the chunk's bounds [cb[2x], cb[2x+1]) for k_rs_downsweep_seg.
template <typename K>
__global__ __launch_bounds__(256) void k_rs_upsweep_seg(const K* __restrict__ in, const uint32_t* __restrict__ totals,
                                                        uint32_t chunk, uint32_t shift, uint32_t* __restrict__ hist2,
                                                        uint32_t* __restrict__ cb) {
  __shared__ uint32_t h[4 * 256];
  __shared__ uint32_t s_a[257], s_b[257], s_tmp[4];
  const int tid = threadIdx.x, wid = tid >> 6;
  for (int i = tid; i < 4 * 256; i += 256) h[i] = 0;
  uint64_t lo, hi;
  (void)seg_chunk<256>(totals, chunk, blockIdx.x, &lo, &hi, s_a, s_b, s_tmp);
  if (tid == 0) {
    cb[2 * (uint64_t)blockIdx.x] = (uint32_t)lo;
    cb[2 * (uint64_t)blockIdx.x + 1] = (uint32_t)hi;
  }
  for (uint64_t t0 = lo; t0 < hi; t0 += SORT_TILE) {
    K k[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint64_t idx = t0 + (uint64_t)i * 256 + tid;
      k[i] = idx < hi ? in[idx] : (K)0;
    }
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint64_t idx = t0 + (uint64_t)i * 256 + tid;
      if (idx < hi) atomicAdd(&h[wid * 256 + ((uint32_t)(k[i] >> shift) & 255u)], 1u);
    }
  }
  __syncthreads();
  hist2[(uint64_t)blockIdx.x * 256 + tid] = h[tid] + h[256 + tid] + h[512 + tid] + h[768 + tid];
}

// Segment s (one workgroup; thread d = digit): the digit totals over the
// segment's chunks, the output base of every (chunk, digit) for
// k_rs_downsweep_seg, and the group table of k_rs_seg.  Group s * 2^b2 + j
// is digit d = (the segment's low 8 - b2 bits) << b2 | j: the keys of a
// segment share the high bits of the digit.  crowd (host-coherent): raised
// when a group exceeds `limit`.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_rs_segbase(const uint32_t* __restrict__ totals, uint32_t chunk,
                                                    const uint32_t* __restrict__ hist2, uint32_t* __restrict__ base2,
                                                    uint32_t b2, uint32_t* __restrict__ gstart,
                                                    uint32_t* __restrict__ gsize, uint32_t* __restrict__ crowd,
                                                    uint32_t limit) {
  __shared__ uint32_t s_a[257], s_b[257], s_tmp[4];
  const uint32_t sg = blockIdx.x, d = threadIdx.x;
  uint64_t c0, c1;
  (void)seg_chunk<256>(totals, chunk, 0u, &c0, &c1, s_a, s_b, s_tmp);  // the tables only
  const uint32_t x0 = s_b[sg], x1 = s_b[sg + 1], start = s_a[sg];
  uint32_t tot = 0;
  for (uint32_t x = x0; x < x1; ++x) tot += hist2[(uint64_t)x * 256 + d];
  const uint32_t dpre = block_excl_scan<256>(tot, s_tmp);
  uint32_t run = start + dpre;
  for (uint32_t x = x0; x < x1; ++x) {
    const uint32_t c = hist2[(uint64_t)x * 256 + d];
    base2[(uint64_t)x * 256 + d] = run;
    run += c;
  }
  const uint32_t lowb = 8u - b2;
  if ((d >> b2) == (sg & ((1u << lowb) - 1u))) {
    const uint32_t g = (sg << b2) | (d & ((1u << b2) - 1u));
    gstart[g] = start + dpre;
    gsize[g] = tot;
    if (tot > limit) __hip_atomic_store(crowd, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The downsweep of the segmented pass: chunk x (blockIdx.x) of its segment,
// with its digit bases from k_rs_segbase.
template <typename K>
__global__ __launch_bounds__((ds_threads<K, false>()), (sizeof(K) == 4 ? 4 : 3)) void k_rs_downsweep_seg(
    const K* __restrict__ in, K* __restrict__ out, const uint32_t* __restrict__ cb, uint32_t shift,
    const uint32_t* __restrict__ base2) {
  constexpr int T = ds_threads<K, false>();
  constexpr int W = T / 64;
  __shared__ __align__(16) K s_keys[T * SORT_ITEMS];
  __shared__ __align__(16) uint32_t s_vals[4];
  __shared__ __align__(16) uint16_t s_cnt[16 * T];
  __shared__ uint32_t s_hist[W * 256];
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_tmp[W];
  const uint32_t tid = threadIdx.x;
  // past the last chunk the range is empty (no early return: it costs the
  // chunk loop 13 spilled registers)
  const uint64_t c0 = cb[2 * (uint64_t)blockIdx.x], c1 = cb[2 * (uint64_t)blockIdx.x + 1];
#pragma unroll
  for (int w = 0; w < W * 256 / T; ++w) s_hist[w * T + tid] = 0;
  if (tid < 256) s_run[tid] = base2[(uint64_t)blockIdx.x * 256 + tid];
  __syncthreads();
  ds_chunk<K, false, false, T>(in, nullptr, out, nullptr, 0u, shift, c0, c1, s_keys, s_vals, s_cnt, s_hist, s_run,
                               s_base, s_tmp);
}

// Group g of a segment sort: [*c0, *c0 + *m) from the table (gstart, gsize),
// or without a table the top-digit segment g of digit_totals (block scan).
template <int T>
__device__ __forceinline__ void seg_group(uint32_t g, const uint32_t* __restrict__ digit_totals,
                                          const uint32_t* __restrict__ gstart, const uint32_t* __restrict__ gsize,
                                          uint32_t* c0, uint32_t* m, uint32_t* s_tmp, uint32_t* s_seg) {
  if (gsize) {
    *c0 = gstart[g];
    *m = gsize[g];
    return;
  }
  const uint32_t c = threadIdx.x < 256 ? digit_totals[threadIdx.x] : 0u;
  const uint32_t pre = block_excl_scan<T>(c, s_tmp);
  if (threadIdx.x == g) {
    s_seg[0] = pre;
    s_seg[1] = c;
  }
  __syncthreads();
  *c0 = s_seg[0];
  *m = s_seg[1];
  __syncthreads();  // s_seg is read before it is written again
}

// A group of m <= SEG_TILE keys sorted in place by its bits [0, nbits) with
// the 4-bit counting passes of the tile sort, back to back in LDS.
template <typename K, int T>
__device__ __forceinline__ void seg_counting_sort(K* __restrict__ seg, uint32_t m, uint32_t nbits, uint32_t* s_keys,
                                                  uint16_t* s_cnt, uint32_t* s_tmp) {
  const uint32_t tid = threadIdx.x;
  const uint32_t lmask = nbits >= 32 ? 0xFFFFFFFFu : (1u << nbits) - 1u;
  const uint32_t v[SORT_ITEMS] = {};  // no values
  const K high = seg[0] & ~(K)lmask;
  uint32_t k[SORT_ITEMS];
  {
    K kk[SORT_ITEMS];
    ts_load_keys<K, false>(seg, 0, m, 0u, kk);
    const uint32_t i0 = tid * SORT_ITEMS;
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) k[i] = i0 + i < m ? (uint32_t)kk[i] & lmask : 0xFFFFFFFFu;
  }
  const uint32_t np = (nbits + 3) / 4;
  for (uint32_t j = 0; j < np; ++j) {
    uint64_t rk;
    ts_count<T>(k, 4 * j, rk, s_cnt);
    __syncthreads();
    (void)ts_scan<T>(s_cnt, 0u, s_tmp);
    __syncthreads();
    ts_scatter<T, uint32_t, false>(k, v, 4 * j, rk, s_cnt, s_keys, nullptr);
    __syncthreads();
    // the next pass writes s_cnt, and s_keys only after its first barrier
    if (j + 1 < np) {
      uint32_t unused[SORT_ITEMS];
      ts_read_blocked<uint32_t, false>(s_keys, nullptr, k, unused);
    }
  }
  for (uint32_t p = tid; p < m; p += T) seg[p] = high | (K)s_keys[ts_swz<4>(p)];
}

// One group per workgroup, sorted in place by its bits [0, nbits) (nbits <=
// 32; the bits above are equal within the group): the keys' top SEG_BB bits
// of the range pick one of SEG_NBK buckets (LDS counters; the scanned starts
// then serve as cursors that place every key in its bucket), and every key's
// place is its bucket's start plus the bucket's smaller keys.  With hashed
// keys a bucket holds ~3 keys.  A group with a bucket of more than
// SEG_RANK_MAX keys (many copies of a few keys) takes the 4-bit counting
// passes instead: in the same workgroup (INLINE: grids of one workgroup per
// CU, where occupancy does not matter), or appended to `slow` (count, then
// group ids) for k_rs_seg_slow, which keeps this kernel at 40 registers and
// four workgroups per CU.  A group above SEG_TILE is never launched here
// (the host sorts crowded keys with the chunked passes) and is left alone.
template <typename K, bool INLINE>
__global__ __launch_bounds__(SEG_THREADS) void k_rs_seg(K* __restrict__ keys, const uint32_t* __restrict__ digit_totals,
                                                        const uint32_t* __restrict__ gstart,
                                                        const uint32_t* __restrict__ gsize, uint32_t nbits,
                                                        uint32_t* __restrict__ slow) {
  constexpr int T = SEG_THREADS, W = T / 64;
  __shared__ __align__(16) uint32_t s_keys[SEG_TILE];
  __shared__ __align__(16) uint16_t s_cnt[INLINE ? 16 * T : 8];  // the counting passes' count table
  __shared__ __align__(16) uint32_t s_c[SEG_NBK / 2];  // bucket counters / cursors, 16 bits each
  __shared__ uint32_t s_tmp[W];
  __shared__ uint32_t s_seg[2];
  const uint32_t tid = threadIdx.x;
  uint32_t c0, m;
  seg_group<T>(blockIdx.x, digit_totals, gstart, gsize, &c0, &m, s_tmp, s_seg);
  if (m <= 1 || m > (uint32_t)SEG_TILE) return;  // the whole workgroup
  K* seg = keys + c0;
  const uint32_t lmask = nbits >= 32 ? 0xFFFFFFFFu : (1u << nbits) - 1u;
  const uint32_t i0 = tid * SORT_ITEMS;
  const uint32_t nreal = m > i0 ? (m - i0 < SORT_ITEMS ? m - i0 : SORT_ITEMS) : 0u;
  const uint32_t bsh = nbits > SEG_BB ? nbits - SEG_BB : 0u;
  uint32_t k[SORT_ITEMS];
  {
    K kk[SORT_ITEMS];
    ts_load_keys<K, false>(seg, 0, m, 0u, kk);
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) k[i] = (uint32_t)kk[i] & lmask;
  }
  const K high = seg[0] & ~(K)lmask;
  for (uint32_t i = tid; i < SEG_NBK / 2; i += T) s_c[i] = 0u;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    if ((uint32_t)i < nreal) {
      const uint32_t b = k[i] >> bsh;
      atomicAdd(&s_c[b >> 1], 1u << (16 * (b & 1u)));
    }
  }
  __syncthreads();
  {
    constexpr uint32_t PT = SEG_NBK / 2 / T;  // counter words per thread
    static_assert(PT % 4 == 0, "whole 16-byte rows of counter words per thread");
    uint4* c4 = reinterpret_cast<uint4*>(s_c) + tid * (PT / 4);
    uint32_t w[PT];
#pragma unroll
    for (uint32_t i = 0; i < PT / 4; ++i) {
      const uint4 a = c4[i];
      w[4 * i] = a.x;
      w[4 * i + 1] = a.y;
      w[4 * i + 2] = a.z;
      w[4 * i + 3] = a.w;
    }
    uint32_t sum = 0, mx = 0;
#pragma unroll
    for (uint32_t i = 0; i < PT; ++i) {
      const uint32_t lo = w[i] & 0xFFFFu, hi = w[i] >> 16;
      sum += lo + hi;
      mx = lo > mx ? lo : mx;
      mx = hi > mx ? hi : mx;
    }
    uint32_t run = block_excl_scan<T>(sum, s_tmp);
    uint32_t st[PT];
#pragma unroll
    for (uint32_t i = 0; i < PT; ++i) {
      const uint32_t lo = run;
      run += w[i] & 0xFFFFu;
      st[i] = lo | (run << 16);
      run += w[i] >> 16;
    }
#pragma unroll
    for (uint32_t i = 0; i < PT / 4; ++i) c4[i] = make_uint4(st[4 * i], st[4 * i + 1], st[4 * i + 2], st[4 * i + 3]);
    if (__syncthreads_or(mx > SEG_RANK_MAX)) {  // also orders the bucket starts
      if (INLINE) seg_counting_sort<K, T>(seg, m, nbits, s_keys, s_cnt, s_tmp);
      else if (tid == 0) slow[1 + atomicAdd(&slow[0], 1u)] = blockIdx.x;
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    if ((uint32_t)i < nreal) {
      const uint32_t b = k[i] >> bsh, hs = 16 * (b & 1u);
      s_keys[(atomicAdd(&s_c[b >> 1], 1u << hs) >> hs) & 0xFFFFu] = k[i];
    }
  }
  __syncthreads();
  // counter b now holds the end of bucket b
  const uint16_t* s_c16 = reinterpret_cast<const uint16_t*>(s_c);
  for (uint32_t p = tid; p < m; p += T) {
    const uint32_t x = s_keys[p], b = x >> bsh;
    const uint32_t st = b ? s_c16[b - 1] : 0u, en = s_c16[b];
    uint32_t rank = st;
    for (uint32_t j = st; j < en; ++j) {
      const uint32_t y = s_keys[j];
      rank += (y < x || (y == x && j < p)) ? 1u : 0u;
    }
    seg[rank] = high | (K)x;  // within the group's lines: merged in L2
  }
}

// The groups k_rs_seg left (slow[0] of them, ids after it): the 4-bit
// counting passes of the tile sort, back to back in LDS.  A fixed grid walks
// the list (empty: every workgroup returns at once).
template <typename K>
__global__ __launch_bounds__(SEG_THREADS) void k_rs_seg_slow(K* __restrict__ keys,
                                                             const uint32_t* __restrict__ digit_totals,
                                                             const uint32_t* __restrict__ gstart,
                                                             const uint32_t* __restrict__ gsize, uint32_t nbits,
                                                             const uint32_t* __restrict__ slow) {
  constexpr int T = SEG_THREADS, W = T / 64;
  __shared__ __align__(16) uint32_t s_keys[SEG_TILE];
  __shared__ __align__(16) uint16_t s_cnt[16 * T];
  __shared__ uint32_t s_tmp[W];
  __shared__ uint32_t s_seg[2];
  const uint32_t ns = slow[0];
  for (uint32_t q = blockIdx.x; q < ns; q += gridDim.x) {
    uint32_t c0, m;
    seg_group<T>(slow[1 + q], digit_totals, gstart, gsize, &c0, &m, s_tmp, s_seg);
    seg_counting_sort<K, T>(keys + c0, m, nbits, s_keys, s_cnt, s_tmp);
    __syncthreads();  // s_keys is read before the next group writes it
  }
}

// Chunks of the segmented pass over n keys (each segment adds at most one
// partial chunk).
inline uint32_t seg_chunks2(uint64_t n, uint32_t chunk) { return (uint32_t)((n + chunk - 1) / chunk) + 256u; }

// seg: room for the segmented pass of a keys-only sort of more than
// 256 * SEG_GROUP_AVG keys (radix_sort with a stream context).
inline size_t sort_workspace_bytes(uint64_t n, size_t key_bytes, bool vals, bool seg = true) {
  const SortPlan p = sort_plan((uint32_t)n);
  size_t b = 2 * aligned_bytes(n * key_bytes) + (vals ? 2 * aligned_bytes(n * 4) : 0) +
             aligned_bytes((size_t)p.G * 256 * 4) + aligned_bytes(256 * 4) + 1024;
  if (seg && !vals) {
    const uint32_t b2 = seg_bits2(n);
    if (b2 >= 1 && b2 <= 8)
      b += 2 * aligned_bytes((size_t)seg_chunks2(n, p.chunk) * 256 * 4) + 2 * aligned_bytes(((size_t)256 << b2) * 4) +
           aligned_bytes((size_t)seg_chunks2(n, p.chunk) * 2 * 4) + aligned_bytes((1 + ((size_t)256 << b2)) * 4);
  }
  return b;
}

// LSD radix sort of n keys over bits [lo_bit, hi_bit).  When hash is set, the
// input is raw uint32 keys, fingerprinted (fp_bits) as they are read in the
// first pass.  vals: nullptr + with_vals -> the original index is carried.
// Returns the buffers holding the result through *keys_out / *vals_out.
// With the stream's context (ds), a keys-only sort partitions on its top
// bits and sorts each group in LDS (segment sort, above); after a crowded
// probe the next 4, 8, ... 256 sorts on the stream skip straight to the
// chunked passes.
template <typename K>
gqf_status radix_sort(hipStream_t s, Arena& ar, const void* in, bool hash, uint32_t fp_bits,
                      const uint32_t* in_vals, bool with_vals, uint32_t n, uint32_t lo_bit, uint32_t hi_bit,
                      K** keys_out, uint32_t** vals_out, DeviceState* ds = nullptr) {
  K* kb[2] = {ar.take<K>(n ? n : 1), ar.take<K>(n ? n : 1)};
  uint32_t* vb[2] = {nullptr, nullptr};
  if (with_vals) {
    vb[0] = ar.take<uint32_t>(n ? n : 1);
    vb[1] = ar.take<uint32_t>(n ? n : 1);
  }
  const SortPlan plan = sort_plan(n);
  uint32_t* hist = ar.take<uint32_t>((size_t)plan.G * 256);
  uint32_t* dbase = ar.take<uint32_t>(256);
  if (ar.used > ar.cap) {
    set_error("internal: sort workspace overflow");
    return GQF_ERR_INVALID;
  }
  if (hi_bit <= lo_bit) hi_bit = lo_bit + 1;
  uint32_t npass = (hi_bit - lo_bit + 7) / 8;
  if (npass > SORT_MAX_PASSES) npass = SORT_MAX_PASSES;
  if (n == 0) {
    *keys_out = kb[0];
    if (vals_out) *vals_out = vb[0];
    return GQF_OK;
  }
  if (n <= (uint32_t)SORT_TILE) {
    // one workgroup, every pass in LDS, one launch
    const double bytes = (double)n * ((hash ? 4.0 : (double)sizeof(K)) + sizeof(K) +
                                      (with_vals ? ((in_vals ? 4.0 : 0.0) + 4.0) : 0.0));
#define GQF_RS_SMALL(V_, H_)                                                                                    \
  GQF_LAUNCH_B(bytes, (k_rs_small<K, V_, H_>), dim3(1), dim3(SORT_THREADS), 0, s, in, in_vals, kb[0], vb[0], n, \
               fp_bits, lo_bit, npass)
    if (with_vals && hash) GQF_RS_SMALL(true, true);
    else if (with_vals) GQF_RS_SMALL(true, false);
    else if (hash) GQF_RS_SMALL(false, true);
    else GQF_RS_SMALL(false, false);
#undef GQF_RS_SMALL
    *keys_out = kb[0];
    if (vals_out) *vals_out = vb[0];
    return GQF_OK;
  }
  const void* src = in;
  const uint32_t* vsrc = in_vals;
  int cur = 0;
  const uint32_t b2 = seg_bits2(n);
  constexpr int cw = sizeof(K) == 8 ? 1 : 0;
  bool probe = ds && !with_vals && lo_bit == 0;
  if (probe && ds->crowd_skip[cw]) {  // recent sorts of this width on this stream were crowded
    --ds->crowd_skip[cw];
    probe = false;
  }
  const bool seg1 = probe && b2 == 0 && hi_bit >= 16 && hi_bit - 8 <= 32;
  const bool seg2 = probe && b2 >= 1 && b2 <= 8 && hi_bit >= 12 + b2 && hi_bit - 8 - b2 <= 32;
  uint32_t *hist2 = nullptr, *base2 = nullptr, *gstart = nullptr, *gsize = nullptr, *cb = nullptr, *slow = nullptr;
  const uint32_t G2 = seg2 ? seg_chunks2(n, plan.chunk) : 0u;
  const uint32_t ngroups = 256u << b2;
  if (seg2) {
    slow = ar.take<uint32_t>(1 + ngroups);
    cb = ar.take<uint32_t>((size_t)G2 * 2);
    hist2 = ar.take<uint32_t>((size_t)G2 * 256);
    base2 = ar.take<uint32_t>((size_t)G2 * 256);
    gstart = ar.take<uint32_t>((size_t)256 << b2);
    gsize = ar.take<uint32_t>((size_t)256 << b2);
    if (ar.used > ar.cap) {
      set_error("internal: sort workspace overflow");
      return GQF_ERR_INVALID;
    }
  }
  if (seg1 || seg2) {
    // the top 8 bits in one chunked pass
    const uint32_t shift = hi_bit - 8;
    *reinterpret_cast<volatile uint32_t*>(ds->h_crowd) = 0u;
    if (hash) {
      GQF_LAUNCH_B((double)n * 4.0, (k_rs_upsweep<K, true>), dim3(plan.G), dim3(256), 0, s, in, n, fp_bits, shift,
                   plan.chunk, plan.G, hist);
    } else {
      GQF_LAUNCH_B((double)n * sizeof(K), (k_rs_upsweep<K, false>), dim3(plan.G), dim3(256), 0, s, in, n, fp_bits,
                   shift, plan.chunk, plan.G, hist);
    }
    GQF_LAUNCH(k_rs_rowscan, dim3(256), dim3(1024), 0, s, hist, plan.G, dbase, seg1 ? ds->d_crowd : (uint32_t*)nullptr,
               (uint32_t)SEG_TILE, slow);
    if (seg1) GQF_HIP(hipEventRecord(ds->ev_crowd, s));
    const double pass_bytes = (double)n * ((hash ? 4.0 : (double)sizeof(K)) + sizeof(K));
    if (hash)
      GQF_LAUNCH_B(pass_bytes, (k_rs_downsweep<K, false, true>), dim3(plan.G), dim3(ds_threads<K, false>()), 0, s, in,
                   nullptr, kb[0], nullptr, n, fp_bits, shift, plan.chunk, plan.G, hist, dbase);
    else
      GQF_LAUNCH_B(pass_bytes, (k_rs_downsweep<K, false, false>), dim3(plan.G), dim3(ds_threads<K, false>()), 0, s, in,
                   nullptr, kb[0], nullptr, n, fp_bits, shift, plan.chunk, plan.G, hist, dbase);
    K* out = kb[0];
    uint32_t nbits = shift;
    if (seg2) {
      // the next b2 bits within every segment
      const uint32_t shift2 = hi_bit - 8 - b2;
      GQF_LAUNCH_B((double)n * sizeof(K), (k_rs_upsweep_seg<K>), dim3(G2), dim3(256), 0, s, kb[0], dbase, plan.chunk,
                   shift2, hist2, cb);
      GQF_LAUNCH(k_rs_segbase, dim3(256), dim3(256), 0, s, dbase, plan.chunk, hist2, base2, b2, gstart, gsize,
                 ds->d_crowd, (uint32_t)SEG_TILE);
      GQF_HIP(hipEventRecord(ds->ev_crowd, s));
      GQF_LAUNCH_B((double)n * 2.0 * sizeof(K), (k_rs_downsweep_seg<K>), dim3(G2), dim3(ds_threads<K, false>()), 0, s,
                   kb[0], kb[1], cb, shift2, base2);
      out = kb[1];
      nbits = shift2;
    }
    // the host reads the flag while the partition's downsweep runs
    GQF_HIP(hipEventSynchronize(ds->ev_crowd));
    const bool crowded = *reinterpret_cast<volatile uint32_t*>(ds->h_crowd) != 0u;
    uint32_t& bo = ds->crowd_backoff[cw];
    bo = crowded ? (bo ? (bo < 256 ? 2 * bo : 256) : 4) : 0;
    ds->crowd_skip[cw] = bo;
    if (!crowded) {
      if (seg2) {
        GQF_LAUNCH_B((double)n * 2.0 * sizeof(K), (k_rs_seg<K, false>), dim3(ngroups), dim3(SEG_THREADS), 0, s, out,
                     (const uint32_t*)nullptr, gstart, gsize, nbits, slow);
        GQF_LAUNCH(k_rs_seg_slow<K>, dim3(SEG_SLOW_GRID), dim3(SEG_THREADS), 0, s, out, (const uint32_t*)nullptr,
                   gstart, gsize, nbits, (const uint32_t*)slow);
      } else {
        GQF_LAUNCH_B((double)n * 2.0 * sizeof(K), (k_rs_seg<K, true>), dim3(256), dim3(SEG_THREADS), 0, s, out, dbase,
                     (const uint32_t*)nullptr, (const uint32_t*)nullptr, nbits, (uint32_t*)nullptr);
      }
      *keys_out = out;
      if (vals_out) *vals_out = nullptr;
      return GQF_OK;
    }
    // a crowded group: every chunked pass over the partitioned fingerprints
    src = out;
    cur = out == kb[0] ? 1 : 0;
    hash = false;
  }
  for (uint32_t p = 0; p < npass; ++p) {
    const uint32_t shift = lo_bit + 8 * p;
    bool first_hash = hash && p == 0;
    // the first pass hashes the raw keys twice, in its upsweep and in its
    // downsweep: with the counting ranks the downsweep has ALU to spare, and
    // that is faster than storing the fingerprints in the upsweep and reading
    // them back (SQF build of 53.7 M keys: 1.17 vs 1.21 ms of kernels)
    if (first_hash) {
      GQF_LAUNCH_B((double)n * 4.0, (k_rs_upsweep<K, true>), dim3(plan.G), dim3(256), 0, s, src, n, fp_bits, shift,
                   plan.chunk, plan.G, hist);
    } else {
      GQF_LAUNCH_B((double)n * sizeof(K), (k_rs_upsweep<K, false>), dim3(plan.G), dim3(256), 0, s, src, n, fp_bits,
                   shift, plan.chunk, plan.G, hist);
    }
    const double kin = first_hash ? 4.0 : (double)sizeof(K);
    GQF_LAUNCH(k_rs_rowscan, dim3(256), dim3(1024), 0, s, hist, plan.G, dbase, (uint32_t*)nullptr, 0u,
               (uint32_t*)nullptr);
    K* dst = kb[cur];
    uint32_t* vdst = vb[cur];
    // read keys (+ values) once, write them once
    const double pass_bytes = (double)n * (kin + sizeof(K) + (with_vals ? ((vsrc ? 4.0 : 0.0) + 4.0) : 0.0));
    if (with_vals) {
      if (first_hash)
        GQF_LAUNCH_B(pass_bytes, (k_rs_downsweep<K, true, true>), dim3(plan.G), dim3(ds_threads<K, true>()), 0, s, src, vsrc,
                     dst, vdst, n, fp_bits, shift, plan.chunk, plan.G, hist, dbase);
      else
        GQF_LAUNCH_B(pass_bytes, (k_rs_downsweep<K, true, false>), dim3(plan.G), dim3(ds_threads<K, true>()), 0, s, src, vsrc,
                     dst, vdst, n, fp_bits, shift, plan.chunk, plan.G, hist, dbase);
    } else {
      if (first_hash)
        GQF_LAUNCH_B(pass_bytes, (k_rs_downsweep<K, false, true>), dim3(plan.G), dim3(ds_threads<K, false>()), 0, s, src,
                     nullptr, dst, nullptr, n, fp_bits, shift, plan.chunk, plan.G, hist, dbase);
      else
        GQF_LAUNCH_B(pass_bytes, (k_rs_downsweep<K, false, false>), dim3(plan.G), dim3(ds_threads<K, false>()), 0, s, src,
                     nullptr, dst, nullptr, n, fp_bits, shift, plan.chunk, plan.G, hist, dbase);
    }
    src = dst;
    vsrc = vdst;
    cur ^= 1;
  }
  *keys_out = (K*)src;
  if (vals_out) *vals_out = (uint32_t*)vsrc;
  return GQF_OK;
}

// ====================================================================== locations
// Canonical location of the i-th kept entry: loc_i = max(fq_i, loc_{i-1}+1).

enum KeepMode { KEEP_ALL = 0, KEEP_UNIQUE = 1, KEEP_TAGGED_INSERT = 2, KEEP_FLAGS = 3 };

// ====================================================================== packed fingerprints
// Read access to an ascending fingerprint array.  Packed form (PK, 64-bit
// values of at most 34 bits): the array holds the low 32 bits, and the high
// word, which only steps up along the array, is the number of the bounds
// h0 <= h1 <= h2 at or below the index (0xFFFFFFFF: a high word not reached).
template <typename K, bool PK>
struct FpArray {
  const void* p;
  uint32_t h0, h1, h2;
  __device__ __forceinline__ K operator[](uint32_t i) const {
    if constexpr (PK) {
      const uint32_t hi = (uint32_t)(i >= h0) + (uint32_t)(i >= h1) + (uint32_t)(i >= h2);
      return (K)(((uint64_t)hi << 32) | reinterpret_cast<const uint32_t*>(p)[i]);
    } else {
      return reinterpret_cast<const K*>(p)[i];
    }
  }
  // the value of entry i from its low word lo (PK)
  __device__ __forceinline__ K widen(uint32_t i, uint32_t lo) const {
    const uint32_t hi = (uint32_t)(i >= h0) + (uint32_t)(i >= h1) + (uint32_t)(i >= h2);
    return (K)(((uint64_t)hi << 32) | lo);
  }
};
// hb: the three bounds in device memory (PK only).
template <typename K, bool PK>
__device__ __forceinline__ FpArray<K, PK> fp_array(const void* p, const uint32_t* __restrict__ hb) {
  FpArray<K, PK> a;
  a.p = p;
  a.h0 = PK ? hb[0] : 0u;
  a.h1 = PK ? hb[1] : 0u;
  a.h2 = PK ? hb[2] : 0u;
  return a;
}

// The first index in [lo, hi) where below(m) fails (below holds, then
// fails), or hi: an 8-ary search by a group of 8 lanes (threadIdx.x % 8 is
// the lane's place in its group).  Each round probes 8 points that cut the
// range into 9 parts, so the chain of dependent loads is log9 of the range
// long instead of log2.  The groups of a wave may search different ranges;
// every lane of the wave calls it (ballots inside).
template <typename P>
__device__ __forceinline__ uint32_t group8_search(uint32_t lo, uint32_t hi, P below) {
  const uint32_t j = threadIdx.x & 7u, sh = threadIdx.x & 56u;  // sh: the group's first lane in the wave
  while (__any(hi - lo >= 9)) {
    const uint64_t span = hi - lo;
    const bool active = span >= 9;
    const bool b = active && below(lo + (uint32_t)(span * (j + 1) / 9));  // a point in (lo, hi)
    const uint32_t c = (uint32_t)__popcll((__ballot(b) >> sh) & 0xFFull);  // probes where below holds
    if (active) {
      const uint32_t nlo = c ? lo + (uint32_t)(span * c / 9) + 1 : lo;
      hi = c < 8 ? lo + (uint32_t)(span * (c + 1) / 9) : hi;
      lo = nlo;
    }
  }
  const bool b = lo + j < hi && below(lo + j);
  return lo + (uint32_t)__popcll((__ballot(b) >> sh) & 0xFFull);
}

// ====================================================================== merge
// Number of A entries among the first d outputs of the stable merge (A first
// on ties).  A and B: pointers or FpArray views.
template <typename AA, typename BB>
__device__ __forceinline__ uint32_t merge_path(AA A, uint32_t na, BB B, uint32_t nb, uint32_t d) {
  uint32_t lo = d > nb ? d - nb : 0u, hi = d < na ? d : na;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (A[mid] <= B[d - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// ====================================================================== delete marking
// keep[i] = 0 for the first c_D(v) stored copies of each value v of A
// (c_D = its multiplicity among the sorted delete fingerprints D).
// Delete marking, multiset difference A - D of two sorted arrays: entry i of A
// is removed iff its rank among the equal entries of A is below the count of
// its value in D (the first cD copies go).  keep[i] = 1 for survivors.
//
// One workgroup per tile of DM_TILE entries of A; the slice of D that can
// match the tile ([lower_bound(first), upper_bound(last)) ) is staged in LDS
// and every thread merges its DM_ITEMS consecutive entries against it.  A
// slice larger than DM_DCAP falls back to per-entry global searches.
constexpr int DM_THREADS = 256;
constexpr int DM_ITEMS = 16;
constexpr int DM_TILE = DM_THREADS * DM_ITEMS;
constexpr uint32_t DM_DCAP = 4096;

template <typename K>
__device__ __forceinline__ uint32_t lower_bound_g(const K* __restrict__ a, uint32_t lo, uint32_t hi, K v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
template <typename K>
__device__ __forceinline__ uint32_t upper_bound_g(const K* __restrict__ a, uint32_t lo, uint32_t hi, K v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The first index in [0, n) whose element is not below x (UPPER: is above
// x), or n, by an 8-ary search over the first 8 lanes of a wave: every lane
// of the wave calls it and gets the result.  Each round probes 8 points
// that cut [lo, hi] into 9 parts, so the chain of dependent loads is
// log9(n) long instead of log2(n).
template <bool UPPER, typename K, typename AA>
__device__ __forceinline__ uint32_t wave8_bound(AA a, uint32_t n, K x) {
  const uint32_t j = threadIdx.x & 63u;
  uint32_t lo = 0, hi = n;  // the answer lies in [lo, hi]
  while (hi - lo >= 9) {
    const uint64_t span = hi - lo;
    bool before = false;  // the probe lies before the answer
    if (j < 8) {
      const K y = a[lo + (uint32_t)(span * (j + 1) / 9)];  // in (lo, hi)
      before = UPPER ? !(x < y) : (y < x);
    }
    const uint32_t c = (uint32_t)__popcll(__ballot(before));  // probes before the answer
    const uint32_t nlo = c ? lo + (uint32_t)(span * c / 9) + 1 : lo;
    hi = c < 8 ? lo + (uint32_t)(span * (c + 1) / 9) : hi;
    lo = nlo;
  }
  bool before = false;
  if (j < 8 && lo + j < hi) {
    const K y = a[lo + j];
    before = UPPER ? !(x < y) : (y < x);
  }
  return lo + (uint32_t)__popcll(__ballot(before));
}

// First index of the equal run of A that contains i (galloping search back).
template <typename AA>
__device__ __forceinline__ uint32_t run_first(AA A, uint32_t i) {
  const auto v = A[i];
  if (i == 0 || A[i - 1] != v) return i;
  uint32_t h = i, step = 1, l;
  for (;;) {
    l = h >= step ? h - step : 0u;
    if (l == 0 || A[l] != v) break;
    h = l;
    step <<= 1;
  }
  if (A[l] == v) return l;
  while (h - l > 1) {
    const uint32_t mid = (l + h) >> 1;
    if (A[mid] == v) h = mid;
    else l = mid;
  }
  return h;
}

// PK: A is packed (FpArray), with its bounds at hb.
template <typename K, bool PK>
__global__ __launch_bounds__(DM_THREADS) void k_delete_mark(const K* __restrict__ Araw, uint32_t na,
                                                            const uint32_t* __restrict__ na_dev,
                                                            const K* __restrict__ D, uint32_t nd,
                                                            uint8_t* __restrict__ keep,
                                                            const uint32_t* __restrict__ hb) {
  __shared__ K s_d[DM_DCAP];
  __shared__ uint32_t s_rng[2];
  const int tid = threadIdx.x;
  const FpArray<K, PK> A = fp_array<K, PK>(Araw, hb);
  if (na_dev) na = *na_dev;  // stored count on the device (the grid covers an upper bound)
  const uint32_t t0 = blockIdx.x * (uint32_t)DM_TILE;
  if (t0 >= na) return;
  const uint32_t t1 = t0 + DM_TILE < na ? t0 + DM_TILE : na;
  // the slice of D this tile can match (waves 0 and 1: uniform branches)
  if (tid < 64) {
    const uint32_t b = wave8_bound<false>(D, nd, A[t0]);
    if (tid == 0) s_rng[0] = b;
  } else if (tid < 128) {
    const uint32_t b = wave8_bound<true>(D, nd, A[t1 - 1]);
    if (tid == 64) s_rng[1] = b;
  }
  // this thread's entries
  const uint32_t i0 = t0 + (uint32_t)tid * DM_ITEMS;
  K v[DM_ITEMS];
  const bool full = i0 + DM_ITEMS <= t1;
  if (full && PK) {
    // 16 low words, 16-byte aligned
    const uint4* p = reinterpret_cast<const uint4*>(reinterpret_cast<const uint32_t*>(Araw) + i0);
#pragma unroll
    for (int k = 0; k < DM_ITEMS / 4; ++k) {
      const uint4 x = p[k];
      v[4 * k] = A.widen(i0 + 4 * k, x.x);
      v[4 * k + 1] = A.widen(i0 + 4 * k + 1, x.y);
      v[4 * k + 2] = A.widen(i0 + 4 * k + 2, x.z);
      v[4 * k + 3] = A.widen(i0 + 4 * k + 3, x.w);
    }
  } else if (full) {
    // A is 256-byte aligned and i0 a multiple of DM_ITEMS: 16-byte aligned
    uint4 w[DM_ITEMS * sizeof(K) / 16];
    const uint4* p = reinterpret_cast<const uint4*>(Araw + i0);
#pragma unroll
    for (int k = 0; k < (int)(DM_ITEMS * sizeof(K) / 16); ++k) w[k] = p[k];
    __builtin_memcpy(v, w, sizeof(v));
  } else {
#pragma unroll
    for (int j = 0; j < DM_ITEMS; ++j) v[j] = i0 + j < t1 ? A[i0 + j] : K(0);
  }
  __syncthreads();
  const uint32_t dlo = s_rng[0], dhi = s_rng[1];
  const uint32_t dn = dhi - dlo;
  const bool staged = dn <= DM_DCAP;
  if (staged)
    for (uint32_t k = tid; k < dn; k += DM_THREADS) s_d[k] = D[dlo + k];
  __syncthreads();
  if (i0 >= t1) return;
  uint32_t out[DM_ITEMS / 4] = {0u, 0u, 0u, 0u};
  // rank of the first entry among its equal run of A
  uint32_t rank = i0 - run_first(A, i0);
  uint32_t p = 0;  // merge position in the D slice
  if (staged) {
    uint32_t lo = 0, hi = dn;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_d[mid] < v[0]) lo = mid + 1;
      else hi = mid;
    }
    p = lo;
  }
#pragma unroll
  for (int j = 0; j < DM_ITEMS; ++j) {
    if (i0 + j < t1) {
      if (j > 0) rank = v[j] == v[j - 1] ? rank + 1 : 0u;
      uint32_t cd;
      if (staged) {
        while (p < dn && s_d[p] < v[j]) ++p;
        uint32_t e = p;
        while (e < dn && s_d[e] == v[j]) ++e;
        cd = e - p;
      } else {
        const uint32_t lo = lower_bound_g(D, dlo, dhi, v[j]);
        cd = upper_bound_g(D, lo, dhi, v[j]) - lo;
      }
      if (rank >= cd) out[j >> 2] |= 1u << (8 * (j & 3));
    }
  }
  if (full) {
    *reinterpret_cast<uint4*>(keep + i0) = make_uint4(out[0], out[1], out[2], out[3]);
  } else {
#pragma unroll
    for (int j = 0; j < DM_ITEMS; ++j)
      if (i0 + j < t1) keep[i0 + j] = (uint8_t)((out[j >> 2] >> (8 * (j & 3))) & 1u);
  }
}

// ====================================================================== region scan
// Exclusive scan of COLS interleaved uint32 columns over `rows` rows, in
// place, by one block of 1024 threads, in pieces of 4096 rows: each thread
// holds 4 consecutive rows in registers (16-byte loads), one block-wide scan
// per column, stores in place; totals are carried between pieces and the
// next piece is loaded before the current one is scanned.  totals[c] =
// column sums (may be NULL).  data must be 16-byte aligned.
template <int COLS>
__device__ __forceinline__ void scan_rows_load(const uint32_t* __restrict__ data, uint32_t rows, uint32_t row0,
                                               uint32_t* v) {
  constexpr int N = 4 * COLS;
  if (row0 + 4 <= rows) {
    const uint4* p = reinterpret_cast<const uint4*>(data + (uint64_t)row0 * COLS);
#pragma unroll
    for (int k = 0; k < N / 4; ++k) {
      const uint4 x = p[k];
      v[4 * k] = x.x;
      v[4 * k + 1] = x.y;
      v[4 * k + 2] = x.z;
      v[4 * k + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = row0 + k / COLS < rows ? data[(uint64_t)row0 * COLS + k] : 0u;
  }
}
constexpr uint32_t SCAN_PIECE = 4096;  // rows per piece (1024 threads x 4 rows)

// part (may be NULL): column sums of every piece (k_scan_rows_sum); workgroup
// g then scans piece g alone on top of the sums of the pieces before it, and
// the last one writes the totals.  Without it one workgroup scans them all.
template <int COLS>
__global__ __launch_bounds__(1024) void k_scan_rows(uint32_t* __restrict__ data, uint32_t rows,
                                                    uint32_t* __restrict__ totals,
                                                    const uint32_t* __restrict__ part) {
  constexpr uint32_t PIECE = SCAN_PIECE;
  constexpr int N = 4 * COLS;
  __shared__ uint32_t s_w[2][COLS][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t base0 = part ? blockIdx.x * PIECE : 0u;
  const uint32_t end = part && base0 + PIECE < rows ? base0 + PIECE : rows;
  uint32_t carry[COLS];
#pragma unroll
  for (int c = 0; c < COLS; ++c) {
    carry[c] = 0;
    if (part)
      for (uint32_t g = 0; g < blockIdx.x; ++g) carry[c] += part[g * COLS + c];
  }
  uint32_t cur[N], nxt[N];
  scan_rows_load<COLS>(data, rows, base0 + (uint32_t)tid * 4, cur);
  int buf = 0;
  for (uint32_t base = base0; base < end; base += PIECE) {
    const uint32_t row0 = base + (uint32_t)tid * 4;
    if (base + PIECE < end) scan_rows_load<COLS>(data, rows, row0 + PIECE, nxt);
    uint32_t pre[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const uint32_t sum = cur[c] + cur[COLS + c] + cur[2 * COLS + c] + cur[3 * COLS + c];
      const uint32_t inc = wave_incl_add(sum);
      if (lane == 63) s_w[buf][c][wid] = inc;
      pre[c] = inc - sum;
    }
    __syncthreads();  // the other buffer of s_w is free again: nobody reads it past this point
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      uint32_t wp = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < 16; ++w) {
        const uint32_t x = s_w[buf][c][w];
        wp += w < wid ? x : 0u;
        tot += x;
      }
      uint32_t x = carry[c] + pre[c] + wp;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t y = cur[j * COLS + c];
        cur[j * COLS + c] = x;
        x += y;
      }
      carry[c] += tot;
    }
    if (row0 + 4 <= rows) {
      uint4* p = reinterpret_cast<uint4*>(data + (uint64_t)row0 * COLS);
#pragma unroll
      for (int k = 0; k < N / 4; ++k) p[k] = make_uint4(cur[4 * k], cur[4 * k + 1], cur[4 * k + 2], cur[4 * k + 3]);
    } else {
#pragma unroll
      for (int k = 0; k < N; ++k)
        if (row0 + k / COLS < rows) data[(uint64_t)row0 * COLS + k] = cur[k];
    }
#pragma unroll
    for (int k = 0; k < N; ++k) cur[k] = nxt[k];
    buf ^= 1;
  }
  if (tid == 0 && totals && (!part || blockIdx.x + 1 == gridDim.x)) {
#pragma unroll
    for (int c = 0; c < COLS; ++c) totals[c] = carry[c];
  }
}

// Column sums of piece blockIdx.x into part[blockIdx.x * COLS ..].
template <int COLS>
__global__ __launch_bounds__(1024) void k_scan_rows_sum(const uint32_t* __restrict__ data, uint32_t rows,
                                                        uint32_t* __restrict__ part) {
  constexpr int N = 4 * COLS;
  __shared__ uint32_t s_w[COLS][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  uint32_t v[N];
  scan_rows_load<COLS>(data, rows, blockIdx.x * SCAN_PIECE + (uint32_t)tid * 4, v);
#pragma unroll
  for (int c = 0; c < COLS; ++c) {
    const uint32_t sum = v[c] + v[COLS + c] + v[2 * COLS + c] + v[3 * COLS + c];
    const uint32_t inc = wave_incl_add(sum);
    if (lane == 63) s_w[c][wid] = inc;
  }
  __syncthreads();
  if (tid < COLS) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += s_w[tid][w];
    part[blockIdx.x * COLS + tid] = t;
  }
}

// Exclusive scan of COLS interleaved columns over `rows` rows in place;
// totals (may be NULL) = column sums.  Several pieces: a sum pass and a scan
// pass of one workgroup per piece (the piece sums from the arena).
template <int COLS>
static inline gqf_status scan_rows(hipStream_t s, Arena& ar, uint32_t* data, uint32_t rows, uint32_t* totals) {
  const uint32_t G = (rows + SCAN_PIECE - 1) / SCAN_PIECE;
  if (G <= 1) {
    GQF_LAUNCH(k_scan_rows<COLS>, dim3(1), dim3(1024), 0, s, data, rows, totals, (const uint32_t*)nullptr);
    return GQF_OK;
  }
  uint32_t* part = ar.take<uint32_t>((size_t)G * COLS);
  if (ar.used > ar.cap) {
    set_error("internal: scan workspace overflow");
    return GQF_ERR_INVALID;
  }
  GQF_LAUNCH(k_scan_rows_sum<COLS>, dim3(G), dim3(1024), 0, s, data, rows, part);
  GQF_LAUNCH(k_scan_rows<COLS>, dim3(G), dim3(1024), 0, s, data, rows, totals, (const uint32_t*)part);
  return GQF_OK;
}
// Workspace bytes of scan_rows.
static inline size_t scan_rows_bytes(uint32_t rows, int cols) {
  return aligned_bytes((size_t)((rows + SCAN_PIECE - 1) / SCAN_PIECE) * cols * 4) + 256;
}

}  // namespace gqf
