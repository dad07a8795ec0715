// gqf_kernels.hpp -- device primitives shared by the SQF and RSQF engines.
//
//   * fingerprint(): Normal_APHash (quotientFilter.cu:134-149 / RSQF.cu:174-189)
//     with the wide extension for q + r > 32;
//   * an LDS-staged LSD radix sort ("onesweep": one read and one write of the
//     keys per 8-bit digit, wave64 ballot ranking, decoupled look-back over
//     tiles), optionally hashing raw keys on the fly and carrying a uint32
//     index -- replaces thrust::sort / sort_by_key (quotientFilter.cu:295,
//     304, 770, 963, 1475; RSQF.cu:433, 442, 913);
//   * the canonical-layout scan: loc_i = max(fq_i, loc_{i-1} + 1) as a
//     single-pass decoupled look-back scan over (count, last location) pairs,
//     fused with duplicate removal / merge rules / delete flags -- replaces the
//     location computations of the three SQF builds (quotientFilter.cu:788-841,
//     985-1022, 1516-1530) and thrust::unique;
//   * merge-path merge with a source tag -- replaces mgpu::merge
//     (quotientFilter.cu:1651, 1800);
//   * delete marking (multiset difference, one stored copy per delete key).
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

__device__ __forceinline__ unsigned long long st_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one uint32 per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block256_excl_scan(uint32_t x, uint32_t* s_tmp4, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
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

// Same for uint64 (used with packed counters).
__device__ __forceinline__ uint64_t block256_excl_scan64(uint64_t x, uint64_t* s_tmp4, uint64_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
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

static __global__ void k_reset_counters(uint32_t* ctr, int from, int to) {
  int i = from + threadIdx.x;
  if (i < to) ctr[i] = 0;
}

template <typename T>
inline void launch_fill(hipStream_t s, T* p, uint64_t n, T v) {
  if (!n) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

// ====================================================================== radix sort
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 4096 keys per tile
constexpr int SORT_MAX_PASSES = 8;
constexpr unsigned long long ST_AGG = 1ull << 30;
constexpr unsigned long long ST_INC = 2ull << 30;
constexpr unsigned long long ST_FLAGS = 3ull << 30;
constexpr unsigned long long ST_VAL = (1ull << 30) - 1ull;

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

template <typename K, bool HASH>
__device__ __forceinline__ K sort_load_key(const void* in, uint64_t idx, uint32_t fp_bits) {
  if (HASH) return (K)fingerprint(reinterpret_cast<const uint32_t*>(in)[idx], fp_bits);
  return reinterpret_cast<const K*>(in)[idx];
}

// Per-pass digit histograms of all keys (one read of the input).
template <typename K, bool HASH>
__global__ __launch_bounds__(256) void k_sort_hist(const void* __restrict__ in, uint32_t n, uint32_t fp_bits,
                                                   uint32_t lo_bit, uint32_t npass, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[SORT_MAX_PASSES * 256];
  for (int i = threadIdx.x; i < SORT_MAX_PASSES * 256; i += 256) h[i] = 0;
  __syncthreads();
  for (uint64_t idx = blockIdx.x * 256ull + threadIdx.x; idx < n; idx += (uint64_t)gridDim.x * 256ull) {
    K key = sort_load_key<K, HASH>(in, idx, fp_bits);
    for (uint32_t p = 0; p < npass; ++p) atomicAdd(&h[p * 256 + ((uint32_t)(key >> (lo_bit + 8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < npass * 256; i += 256) {
    uint32_t c = h[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// Exclusive scan of each pass's 256 digit counts, in place (grid = npass).
static __global__ __launch_bounds__(256) void k_sort_hist_scan(uint32_t* hist) {
  __shared__ uint32_t tmp[4];
  uint32_t* hp = hist + blockIdx.x * 256;
  uint32_t x = hp[threadIdx.x];
  uint32_t e = block256_excl_scan(x, tmp, nullptr);
  hp[threadIdx.x] = e;
}

// One digit pass.  Tiles are claimed in launch order through an atomic
// counter so every tile a look-back waits on is already running.
template <typename K, bool VALS, bool HASH>
__global__ __launch_bounds__(256) void k_onesweep(const void* __restrict__ in_keys, const uint32_t* __restrict__ in_vals,
                                                  K* __restrict__ out_keys, uint32_t* __restrict__ out_vals,
                                                  uint32_t n, uint32_t shift, uint32_t fp_bits,
                                                  const uint32_t* __restrict__ digit_base,
                                                  unsigned long long* __restrict__ status,
                                                  uint32_t* __restrict__ tile_counter, uint32_t epoch) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wcnt[4 * 256];
  __shared__ uint32_t s_lstart[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_tmp[4];
  __shared__ K s_keys[SORT_TILE];
  __shared__ uint32_t s_vals[VALS ? SORT_TILE : 1];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = tid; i < 4 * 256; i += 256) s_wcnt[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t base = (uint64_t)tile * SORT_TILE + (uint64_t)wid * (SORT_ITEMS * 64) + lane;

  K k[SORT_ITEMS];
  uint32_t v[VALS ? SORT_ITEMS : 1];
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint64_t idx = base + (uint64_t)i * 64;
    const bool valid = idx < n;
    // padding keys (all ones) take the last digit and rank after real keys
    k[i] = valid ? sort_load_key<K, HASH>(in_keys, idx, fp_bits) : (K)~(K)0;
    if (VALS) v[i] = in_vals ? (valid ? in_vals[idx] : 0u) : (uint32_t)idx;
  }

  // stable rank within the wave: items are visited in index order
  uint32_t rk[SORT_ITEMS];
  const uint64_t lt = (1ull << lane) - 1ull;
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

  // thread d owns digit d: wave prefix, tile total, look-back
  {
    const uint32_t d = tid;
    const uint32_t c0 = s_wcnt[d], c1 = s_wcnt[256 + d], c2 = s_wcnt[512 + d], c3 = s_wcnt[768 + d];
    const uint32_t total = c0 + c1 + c2 + c3;
    const unsigned long long tag = (unsigned long long)epoch << 32;
    unsigned long long* my = status + (uint64_t)tile * 256 + d;
    st_store(my, tag | (tile == 0 ? ST_INC : ST_AGG) | total);
    s_wcnt[d] = 0;
    s_wcnt[256 + d] = c0;
    s_wcnt[512 + d] = c0 + c1;
    s_wcnt[768 + d] = c0 + c1 + c2;
    s_lstart[d] = block256_excl_scan(total, s_tmp, nullptr);
    uint32_t excl = 0;
    if (tile > 0) {
      uint32_t j = tile - 1;
      for (;;) {
        unsigned long long st;
        do {
          st = st_load(status + (uint64_t)j * 256 + d);
        } while ((uint32_t)(st >> 32) != epoch || (st & ST_FLAGS) == 0);
        excl += (uint32_t)(st & ST_VAL);
        if ((st & ST_FLAGS) == ST_INC) break;
        --j;
      }
      st_store(my, tag | ST_INC | (unsigned long long)(excl + total));
    }
    s_gbase[d] = digit_base[d] + excl;
  }
  __syncthreads();

#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
    const uint32_t p = s_lstart[d] + s_wcnt[wid * 256 + d] + rk[i];
    s_keys[p] = k[i];
    if (VALS) s_vals[p] = v[i];
  }
  __syncthreads();

  const uint64_t left = (uint64_t)n - (uint64_t)tile * SORT_TILE;
  const uint32_t tile_n = left < (uint64_t)SORT_TILE ? (uint32_t)left : (uint32_t)SORT_TILE;
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    const uint32_t p = tid + i * SORT_THREADS;
    if (p < tile_n) {
      const K key = s_keys[p];
      const uint32_t d = (uint32_t)(key >> shift) & 255u;
      const uint32_t g = s_gbase[d] + (p - s_lstart[d]);
      out_keys[g] = key;
      if (VALS) out_vals[g] = s_vals[p];
    }
  }
}

struct SortBuffers {
  void* keys[2];
  uint32_t* vals[2];
  uint32_t* hist;
};

inline size_t sort_workspace_bytes(uint64_t n, size_t key_bytes, bool vals) {
  return 2 * aligned_bytes(n * key_bytes) + (vals ? 2 * aligned_bytes(n * 4) : 0) +
         aligned_bytes(SORT_MAX_PASSES * 256 * 4) + 1024;
}
inline size_t sort_status_words(uint64_t n) { return ((n + SORT_TILE - 1) / SORT_TILE) * 256 + 256; }

// LSD radix sort of n keys over bits [lo_bit, hi_bit).  When hash is set, the
// input is raw uint32 keys, fingerprinted (fp_bits) as they are first read.
// vals: nullptr + with_vals -> the original index is carried.  Returns the
// buffers holding the result through *keys_out / *vals_out.
template <typename K>
gqf_status radix_sort(hipStream_t s, DeviceState& ds, Arena& ar, const void* in, bool hash, uint32_t fp_bits,
                      const uint32_t* in_vals, bool with_vals, uint32_t n, uint32_t lo_bit, uint32_t hi_bit,
                      K** keys_out, uint32_t** vals_out) {
  K* kb[2] = {ar.take<K>(n ? n : 1), ar.take<K>(n ? n : 1)};
  uint32_t* vb[2] = {nullptr, nullptr};
  if (with_vals) {
    vb[0] = ar.take<uint32_t>(n ? n : 1);
    vb[1] = ar.take<uint32_t>(n ? n : 1);
  }
  uint32_t* hist = ar.take<uint32_t>(SORT_MAX_PASSES * 256);
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
  const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
  GQF_TRY(status_reserve(ds, (size_t)tiles * 256 + 256));
  launch_fill<uint32_t>(s, hist, SORT_MAX_PASSES * 256, 0u);
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, ds.counters, (int)CTR_SORT_TILE,
                     (int)CTR_SORT_TILE + SORT_MAX_PASSES);
  {
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (hash)
      hipLaunchKernelGGL((k_sort_hist<K, true>), dim3(blocks), dim3(256), 0, s, in, n, fp_bits, lo_bit, npass, hist);
    else
      hipLaunchKernelGGL((k_sort_hist<K, false>), dim3(blocks), dim3(256), 0, s, in, n, fp_bits, lo_bit, npass, hist);
  }
  hipLaunchKernelGGL(k_sort_hist_scan, dim3(npass), dim3(256), 0, s, hist);
  const void* src = in;
  const uint32_t* vsrc = in_vals;
  int cur = 0;
  for (uint32_t p = 0; p < npass; ++p) {
    const uint32_t shift = lo_bit + 8 * p;
    const uint32_t epoch = ++ds.epoch;
    K* dst = kb[cur];
    uint32_t* vdst = vb[cur];
    const bool first_hash = hash && p == 0;
    if (with_vals) {
      if (first_hash)
        hipLaunchKernelGGL((k_onesweep<K, true, true>), dim3(tiles), dim3(SORT_THREADS), 0, s, src, vsrc, dst, vdst, n,
                           shift, fp_bits, hist + p * 256, ds.status, ds.counters + CTR_SORT_TILE + p, epoch);
      else
        hipLaunchKernelGGL((k_onesweep<K, true, false>), dim3(tiles), dim3(SORT_THREADS), 0, s, src, vsrc, dst, vdst,
                           n, shift, fp_bits, hist + p * 256, ds.status, ds.counters + CTR_SORT_TILE + p, epoch);
    } else {
      if (first_hash)
        hipLaunchKernelGGL((k_onesweep<K, false, true>), dim3(tiles), dim3(SORT_THREADS), 0, s, src, nullptr, dst,
                           nullptr, n, shift, fp_bits, hist + p * 256, ds.status, ds.counters + CTR_SORT_TILE + p, epoch);
      else
        hipLaunchKernelGGL((k_onesweep<K, false, false>), dim3(tiles), dim3(SORT_THREADS), 0, s, src, nullptr, dst,
                           nullptr, n, shift, fp_bits, hist + p * 256, ds.status, ds.counters + CTR_SORT_TILE + p,
                           epoch);
    }
    src = dst;
    vsrc = vdst;
    cur ^= 1;
  }
  *keys_out = (K*)src;
  if (vals_out) *vals_out = (uint32_t*)vsrc;
  return GQF_OK;
}

// ====================================================================== layout scan
// Canonical location of the i-th kept entry: loc_i = max(fq_i, loc_{i-1}+1).
// A block of entries acts on the incoming location x as x -> max(x + c, l)
// where c = entries kept and l = last location when nothing precedes; the
// pair (c, l) composes associatively: (a then b) = (ca+cb, max(la+cb, lb)).
constexpr int LS_THREADS = 256;
constexpr int LS_ITEMS = 16;
constexpr int LS_TILE = LS_THREADS * LS_ITEMS;
constexpr int64_t LOC_NONE = -(1ll << 40);

enum KeepMode { KEEP_ALL = 0, KEEP_UNIQUE = 1, KEEP_TAGGED_INSERT = 2, KEEP_FLAGS = 3 };

struct CL {
  uint32_t c;
  int64_t l;
};
__device__ __forceinline__ CL cl_combine(CL a, CL b) {
  CL o;
  o.c = a.c + b.c;
  const int64_t x = a.l == LOC_NONE ? LOC_NONE : a.l + (int64_t)b.c;
  o.l = x > b.l ? x : b.l;
  return o;
}
__device__ __forceinline__ unsigned long long cl_pack(unsigned long long flag, CL x) {
  unsigned long long lp = 0;
  if (x.l != LOC_NONE) lp = (unsigned long long)(x.l + 1 < 0xFFFFFFFFll ? x.l + 1 : 0xFFFFFFFFll);
  return (flag << 62) | ((unsigned long long)x.c << 32) | lp;
}
__device__ __forceinline__ CL cl_unpack(unsigned long long w) {
  CL o;
  o.c = (uint32_t)((w >> 32) & 0x3FFFFFFFull);
  const uint32_t lp = (uint32_t)w;
  o.l = lp == 0 ? LOC_NONE : (int64_t)lp - 1;
  return o;
}

// in: sorted entries (fingerprint << tagbits | tag).  out: the kept
// fingerprints (ufp) and their canonical locations (loc), compacted; the kept
// count goes to ctr[CTR_COUNT]; overflow of capacity (or, for the RSQF, of
// the 8-bit block offset) raises DEV_ERR_CAPACITY in ctr[CTR_ERR].
template <typename K, int MODE>
__global__ __launch_bounds__(256) void k_layout_scan(const K* __restrict__ in, const uint8_t* __restrict__ flags,
                                                     uint32_t n, uint32_t q, uint32_t r, uint32_t tagbits,
                                                     uint64_t capacity, int rsqf_offsets, K* __restrict__ ufp,
                                                     uint32_t* __restrict__ loc, unsigned long long* __restrict__ status,
                                                     uint32_t* __restrict__ ctr) {
  __shared__ uint32_t s_tile;
  __shared__ K s_buf[LS_TILE + 1];
  __shared__ uint32_t s_loc[LS_TILE];
  __shared__ uint32_t s_wc[4];
  __shared__ int64_t s_wl[4];
  __shared__ uint32_t s_pc;
  __shared__ int64_t s_pl;
  __shared__ uint32_t s_tc;

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(&ctr[CTR_SCAN_TILE], 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t t0 = (uint64_t)tile * LS_TILE;
  const uint32_t tn = (uint32_t)((uint64_t)n - t0 < (uint64_t)LS_TILE ? (uint64_t)n - t0 : (uint64_t)LS_TILE);
  for (uint32_t i = tid; i < tn; i += LS_THREADS) s_buf[i + 1] = in[t0 + i];
  if (tid == 0) s_buf[0] = t0 > 0 ? in[t0 - 1] : (K)0;
  __syncthreads();

  const uint64_t mq = mask64(q);
  K x[LS_ITEMS];
  uint32_t keepbits = 0;
  CL agg = {0u, LOC_NONE};
  const uint32_t i0 = (uint32_t)tid * LS_ITEMS;
#pragma unroll
  for (int j = 0; j < LS_ITEMS; ++j) {
    const uint32_t i = i0 + j;
    x[j] = 0;
    if (i < tn) {
      const K key = s_buf[i + 1];
      const K fp = key >> tagbits;
      const bool first = (t0 + i) == 0;
      bool kp;
      if (MODE == KEEP_ALL) kp = true;
      else if (MODE == KEEP_UNIQUE) kp = first || fp != (s_buf[i] >> tagbits);
      else if (MODE == KEEP_TAGGED_INSERT) kp = ((key & (K)1) == 0) || first || fp != (s_buf[i] >> 1);
      else kp = flags[t0 + i] != 0;
      x[j] = fp;
      if (kp) {
        keepbits |= 1u << j;
        const int64_t fq = (int64_t)((uint64_t)(fp >> r) & mq);
        agg.l = agg.l == LOC_NONE ? fq : (fq > agg.l + 1 ? fq : agg.l + 1);
        agg.c++;
      }
    }
  }

  // block-wide exclusive scan of the per-thread (c, l)
  CL inc = agg;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    CL y;
    y.c = __shfl_up(inc.c, d, 64);
    y.l = __shfl_up(inc.l, d, 64);
    if (lane >= d) inc = cl_combine(y, inc);
  }
  CL ex;
  ex.c = __shfl_up(inc.c, 1, 64);
  ex.l = __shfl_up(inc.l, 1, 64);
  if (lane == 0) {
    ex.c = 0;
    ex.l = LOC_NONE;
  }
  if (lane == 63) {
    s_wc[wid] = inc.c;
    s_wl[wid] = inc.l;
  }
  __syncthreads();
  CL wpre = {0u, LOC_NONE}, tagg = {0u, LOC_NONE};
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    CL wv = {s_wc[w], s_wl[w]};
    if (w < wid) wpre = cl_combine(wpre, wv);
    tagg = cl_combine(tagg, wv);
  }
  ex = cl_combine(wpre, ex);

  if (tid == 0) {
    CL excl = {0u, LOC_NONE};
    unsigned long long* my = status + tile;
    if (tile == 0) {
      st_store(my, cl_pack(2, tagg));
    } else {
      st_store(my, cl_pack(1, tagg));
      uint32_t j = tile - 1;
      for (;;) {
        unsigned long long st;
        do {
          st = st_load(status + j);
        } while ((st >> 62) == 0);
        excl = cl_combine(cl_unpack(st), excl);
        if ((st >> 62) == 2) break;
        --j;
      }
      st_store(my, cl_pack(2, cl_combine(excl, tagg)));
    }
    s_pc = excl.c;
    s_pl = excl.l;
    s_tc = tagg.c;
    if (tile == gridDim.x - 1) {
      const CL tot = cl_combine(excl, tagg);
      ctr[CTR_COUNT] = tot.c;
      if (tot.l != LOC_NONE && (uint64_t)tot.l >= capacity) atomicOr(&ctr[CTR_ERR], DEV_ERR_CAPACITY);
    }
  }
  __syncthreads();

  // emit kept entries with their locations into LDS, then write coalesced
  const CL pre = {s_pc, s_pl};
  const CL me = cl_combine(pre, ex);
  int64_t prev = me.l;
  uint32_t o = me.c - s_pc;
  bool off_overflow = false;
#pragma unroll
  for (int j = 0; j < LS_ITEMS; ++j) {
    if (keepbits & (1u << j)) {
      const int64_t fq = (int64_t)((uint64_t)(x[j] >> r) & mq);
      const int64_t lc = prev == LOC_NONE ? fq : (fq > prev + 1 ? fq : prev + 1);
      if (rsqf_offsets && lc + 1 - 64 * (fq / 64 + 1) > 255) off_overflow = true;
      s_buf[o] = x[j];
      s_loc[o] = (uint32_t)lc;
      prev = lc;
      ++o;
    }
  }
  if (off_overflow) atomicOr(&ctr[CTR_ERR], DEV_ERR_CAPACITY);
  __syncthreads();
  const uint32_t tc = s_tc, pc = s_pc;
  for (uint32_t i = tid; i < tc; i += LS_THREADS) {
    ufp[pc + i] = s_buf[i];
    loc[pc + i] = s_loc[i];
  }
}

// Returns the layout outputs; kept count lands in ds.counters[CTR_COUNT].
template <typename K, int MODE>
gqf_status layout_scan(hipStream_t s, DeviceState& ds, Arena& ar, const K* in, const uint8_t* flags, uint32_t n,
                       uint32_t q, uint32_t r, uint32_t tagbits, uint64_t capacity, bool rsqf_offsets, K** ufp_out,
                       uint32_t** loc_out) {
  const uint32_t tiles = (n + LS_TILE - 1) / LS_TILE;
  K* ufp = ar.take<K>(n ? n : 1);
  uint32_t* loc = ar.take<uint32_t>(n ? n : 1);
  unsigned long long* st = ar.take<unsigned long long>(tiles + 1);
  if (ar.used > ar.cap) {
    set_error("internal: layout workspace overflow");
    return GQF_ERR_INVALID;
  }
  *ufp_out = ufp;
  *loc_out = loc;
  if (n == 0) return GQF_OK;  // CTR_COUNT stays 0 (reset by the caller)
  launch_fill<unsigned long long>(s, st, tiles + 1, 0ull);
  hipLaunchKernelGGL((k_layout_scan<K, MODE>), dim3(tiles), dim3(LS_THREADS), 0, s, in, flags, n, q, r, tagbits,
                     capacity, (int)rsqf_offsets, ufp, loc, st, ds.counters);
  return GQF_OK;
}
inline size_t layout_workspace_bytes(uint64_t n, size_t key_bytes) {
  return aligned_bytes(n * key_bytes) + aligned_bytes(n * 4) + aligned_bytes(((n + LS_TILE - 1) / LS_TILE + 1) * 8) +
         768;
}

// ====================================================================== merge
constexpr int MG_THREADS = 256;
constexpr int MG_ITEMS = 8;
constexpr int MG_TILE = MG_THREADS * MG_ITEMS;

// Number of A entries among the first d outputs of the stable merge (A first on ties).
template <typename K>
__device__ __forceinline__ uint32_t merge_path(const K* A, uint32_t na, const K* B, uint32_t nb, uint32_t d) {
  uint32_t lo = d > nb ? d - nb : 0u, hi = d < na ? d : na;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (A[mid] <= B[d - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// out = merge(A, B) as (fp << 1 | source): A entries carry tag 0 and precede
// equal B entries (tag 1).
template <typename K>
__global__ __launch_bounds__(256) void k_merge_tagged(const K* __restrict__ A, uint32_t na, const K* __restrict__ B,
                                                      uint32_t nb, uint64_t* __restrict__ out) {
  __shared__ K sA[MG_TILE];
  __shared__ K sB[MG_TILE];
  __shared__ uint32_t s_split[2];
  const uint32_t tid = threadIdx.x;
  const uint64_t total = (uint64_t)na + nb;
  const uint32_t d0 = blockIdx.x * MG_TILE;
  const uint32_t d1 = (uint32_t)((uint64_t)d0 + MG_TILE < total ? (uint64_t)d0 + MG_TILE : total);
  if (tid == 0) s_split[0] = merge_path(A, na, B, nb, d0);
  if (tid == 1) s_split[1] = merge_path(A, na, B, nb, d1);
  __syncthreads();
  const uint32_t a0 = s_split[0], a1 = s_split[1];
  const uint32_t b0 = d0 - a0, b1 = d1 - a1;
  const uint32_t la = a1 - a0, lb = b1 - b0;
  for (uint32_t i = tid; i < la; i += MG_THREADS) sA[i] = A[a0 + i];
  for (uint32_t i = tid; i < lb; i += MG_THREADS) sB[i] = B[b0 + i];
  __syncthreads();
  const uint32_t len = la + lb;
  const uint32_t ld = tid * MG_ITEMS < len ? tid * MG_ITEMS : len;
  uint32_t ai = merge_path(sA, la, sB, lb, ld);
  uint32_t bi = ld - ai;
#pragma unroll
  for (int j = 0; j < MG_ITEMS; ++j) {
    const uint32_t o = ld + j;
    if (o < len) {
      const bool takeA = bi >= lb || (ai < la && sA[ai] <= sB[bi]);
      out[(uint64_t)d0 + o] = takeA ? ((uint64_t)sA[ai++] << 1) : (((uint64_t)sB[bi++] << 1) | 1ull);
    }
  }
}

template <typename K>
inline void launch_merge_tagged(hipStream_t s, const K* A, uint32_t na, const K* B, uint32_t nb, uint64_t* out) {
  const uint64_t total = (uint64_t)na + nb;
  if (!total) return;
  hipLaunchKernelGGL(k_merge_tagged<K>, dim3((unsigned)((total + MG_TILE - 1) / MG_TILE)), dim3(MG_THREADS), 0, s, A,
                     na, B, nb, out);
}

// ====================================================================== delete marking
// keep[i] = 0 for the first c_D(v) stored copies of each value v of A
// (c_D = its multiplicity among the sorted delete fingerprints D).
template <typename K>
__global__ __launch_bounds__(256) void k_delete_mark(const K* __restrict__ A, uint32_t na, const K* __restrict__ D,
                                                     uint32_t nd, uint8_t* __restrict__ keep) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= na) return;
  const K v = A[i];
  uint32_t lo = 0, hi = nd;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (D[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  uint32_t ub_lo = lo, ub_hi = nd;
  while (ub_lo < ub_hi) {
    const uint32_t mid = (ub_lo + ub_hi) >> 1;
    if (D[mid] <= v) ub_lo = mid + 1;
    else ub_hi = mid;
  }
  const uint32_t cd = ub_lo - lo;
  if (cd == 0) {
    keep[i] = 1;
    return;
  }
  // rank of entry i among equal entries of A (galloping search backwards)
  uint32_t first = (uint32_t)i;
  if (i > 0 && A[i - 1] == v) {
    uint32_t h = (uint32_t)i, step = 1, l;
    for (;;) {
      l = h >= step ? h - step : 0u;
      if (l == 0 || A[l] != v) break;
      h = l;
      step <<= 1;
    }
    if (A[l] == v) {
      first = l;
    } else {
      while (h - l > 1) {
        const uint32_t mid = (l + h) >> 1;
        if (A[mid] == v) h = mid;
        else l = mid;
      }
      first = h;
    }
  }
  keep[i] = ((uint32_t)i - first) >= cd ? 1 : 0;
}

// ====================================================================== region scan
// Exclusive scan of `cols` interleaved uint32 columns over `rows` rows, in
// place, by one block of 1024 threads (rows are region counts: small).
static __global__ __launch_bounds__(1024) void k_scan_rows(uint32_t* __restrict__ data, uint32_t rows, uint32_t cols,
                                                    uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (uint32_t c = 0; c < cols; ++c) {
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < rows; b += 1024) {
      const uint32_t i = b + tid;
      const uint32_t x = i < rows ? data[(uint64_t)i * cols + c] : 0u;
      uint32_t inc = x;
      for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      if (lane == 63) s_w[wid] = inc;
      __syncthreads();
      uint32_t pre = 0, tot = 0;
      for (int w = 0; w < 16; ++w) {
        pre += w < wid ? s_w[w] : 0u;
        tot += s_w[w];
      }
      const uint32_t carry = s_carry;
      if (i < rows) data[(uint64_t)i * cols + c] = carry + pre + inc - x;
      __syncthreads();
      if (tid == 0) s_carry = carry + tot;
      __syncthreads();
    }
    if (tid == 0 && totals) totals[c] = s_carry;
    __syncthreads();
  }
}

}  // namespace gqf
