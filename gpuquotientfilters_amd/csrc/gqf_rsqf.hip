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
#include "gqf_kernels.hpp"

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

template <int R>
__global__ __launch_bounds__(256) void k_rsqf_lookup(RsqfDev f, uint32_t q, const uint32_t* __restrict__ keys,
                                                     uint32_t n, uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint64_t fp = fingerprint(keys[i], q + f.r);
  out[i] = rsqf_lookup_one<R>(f, (uint32_t)((fp >> f.r) & mask64(q)), (uint32_t)(fp & ((1u << f.r) - 1u)));
}

template <int R, typename K>
__global__ __launch_bounds__(256) void k_rsqf_lookup_sorted(RsqfDev f, uint32_t q, const K* __restrict__ fps,
                                                            const uint32_t* __restrict__ idx, uint32_t n,
                                                            uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint64_t fp = (uint64_t)fps[i];
  out[idx[i]] = rsqf_lookup_one<R>(f, (uint32_t)((fp >> f.r) & mask64(q)), (uint32_t)(fp & ((1u << f.r) - 1u)));
}

// ---------------------------------------------------------------- table writer
// One workgroup per 64 blocks: occupieds / runEnds / remainders of the
// canonical layout assembled in LDS, offsets from the last entry whose
// quotient precedes each block, then every byte of the group written once.
constexpr uint32_t RW_BLOCKS = 64;

template <typename K>
__device__ __forceinline__ uint32_t lb_loc(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint64_t v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
template <typename K>
__device__ __forceinline__ uint32_t lb_fq(const K* __restrict__ fp, uint32_t lo, uint32_t hi, uint64_t v, uint32_t r) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (((uint64_t)fp[mid] >> r) < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K>
__global__ __launch_bounds__(256) void k_rsqf_write(uint8_t* __restrict__ base, uint32_t nb, uint32_t r,
                                                    const K* __restrict__ ufp, const uint32_t* __restrict__ loc,
                                                    const uint32_t* __restrict__ ctr) {
  if (ctr[CTR_ERR]) return;
  __shared__ unsigned long long s_occ[RW_BLOCKS];
  __shared__ unsigned long long s_ends[RW_BLOCKS];
  __shared__ uint32_t s_off[RW_BLOCKS];
  __shared__ uint32_t s_rem[RW_BLOCKS * 16 * 2];  // 8*r bytes per block, r <= 16
  __shared__ uint32_t s_rng[4];
  const uint32_t m = ctr[CTR_COUNT];
  const uint32_t tid = threadIdx.x;
  const uint32_t j0 = blockIdx.x * RW_BLOCKS;
  const uint32_t j1 = j0 + RW_BLOCKS < nb ? j0 + RW_BLOCKS : nb;
  const uint64_t sl0 = (uint64_t)j0 * 64, sl1 = (uint64_t)j1 * 64;
  const uint32_t remw = (j1 - j0) * 2 * r;  // 32-bit words of remainder bits in this group
  for (uint32_t i = tid; i < RW_BLOCKS; i += 256) {
    s_occ[i] = 0;
    s_ends[i] = 0;
    s_off[i] = 0;
  }
  for (uint32_t i = tid; i < remw; i += 256) s_rem[i] = 0;
  if (tid == 0) s_rng[0] = lb_loc<K>(loc, 0u, m, sl0);
  if (tid == 64) s_rng[1] = lb_loc<K>(loc, 0u, m, sl1);
  if (tid == 128) s_rng[2] = lb_fq(ufp, 0u, m, sl0, r);
  if (tid == 192) s_rng[3] = lb_fq(ufp, 0u, m, sl1, r);
  __syncthreads();
  const uint32_t f0 = s_rng[2], f1 = s_rng[3];
  for (uint32_t i = s_rng[0] + tid; i < s_rng[1]; i += 256) {
    const uint64_t fp = (uint64_t)ufp[i];
    const uint64_t fq = fp >> r;
    const uint32_t ls = (uint32_t)(loc[i] - sl0);  // slot within the group
    const uint32_t fr = (uint32_t)(fp & ((1ull << r) - 1ull));
    const uint32_t bit = ls * r;  // 64*r bits per block: group-relative bit position
    atomicOr(&s_rem[bit >> 5], fr << (bit & 31));
    if ((bit & 31) + r > 32) atomicOr(&s_rem[(bit >> 5) + 1], fr >> (32 - (bit & 31)));
    const bool run_end = (i + 1 == m) || (((uint64_t)ufp[i + 1] >> r) != fq);
    if (run_end) atomicOr(&s_ends[ls >> 6], 1ull << (ls & 63));
  }
  for (uint32_t i = f0 + tid; i < f1; i += 256) {
    const uint64_t fq = ((uint64_t)ufp[i] >> r) - sl0;
    atomicOr(&s_occ[fq >> 6], 1ull << (fq & 63));
  }
  // offset_j: entries with quotient < 64 j that sit in block j or later
  for (uint32_t b = tid; b < j1 - j0; b += 256) {
    const uint64_t lim = (uint64_t)(j0 + b) * 64;
    const uint32_t lo = f0 > 0 ? f0 - 1 : 0u;
    const uint32_t idx = lb_fq(ufp, lo, f1, lim, r);
    if (idx > 0) {
      const uint64_t E = loc[idx - 1];
      if (E >= lim) s_off[b] = (uint32_t)(E - lim + 1) & 0xFFu;
    }
  }
  __syncthreads();
  uint64_t* meta = reinterpret_cast<uint64_t*>(base);
  for (uint32_t i = tid; i < (j1 - j0) * 4; i += 256) {
    const uint32_t b = i >> 2, w = i & 3;
    const uint64_t v = w == 0 ? s_occ[b] : w == 1 ? s_ends[b] : w == 2 ? (uint64_t)s_off[b] : 0ull;
    meta[(size_t)(j0 + b) * 4 + w] = v;
  }
  uint32_t* remg = reinterpret_cast<uint32_t*>(base + rsqf_meta_bytes(nb) + (size_t)j0 * 8 * r);
  for (uint32_t i = tid; i < remw; i += 256) remg[i] = s_rem[i];
}

// ---------------------------------------------------------------- decode
// Block j owns the runs of its occupied quotients; the first of them starts at
// slot 64 j + offset_j (all earlier quotients' entries end before it), and the
// k-th owned run ends at the k-th run end from there on.  Requires a canonical
// table (no gaps between a run's predecessor and its quotient's first slot).
__device__ __forceinline__ uint64_t next_run_end(const RsqfDev& f, uint64_t pos) {
  uint32_t b = (uint32_t)(pos >> 6);
  if (b >= f.nb) return ~0ull;
  uint64_t w = f.ends(b) & ~low_mask_ll((uint32_t)(pos & 63));
  while (w == 0) {
    if (++b >= f.nb) return ~0ull;
    w = f.ends(b);
  }
  return (uint64_t)b * 64 + (uint64_t)__ffsll((long long)w) - 1;
}

template <bool EMIT, typename K>
__global__ __launch_bounds__(256) void k_rsqf_decode(RsqfDev f, uint32_t* __restrict__ counts,
                                                     K* __restrict__ out) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= f.nb) return;
  uint64_t occw = f.occ(j);
  uint64_t pos = (uint64_t)j * 64 + f.offset(j);
  int64_t prevEnd = (int64_t)pos - 1;
  uint32_t o = EMIT ? counts[j] : 0u;
  uint32_t cnt = 0;
  while (occw) {
    const uint32_t bit = (uint32_t)__ffsll((long long)occw) - 1;
    occw &= occw - 1;
    const uint64_t fq = (uint64_t)j * 64 + bit;
    const uint64_t e = next_run_end(f, pos);
    if (e == ~0ull) break;
    const int64_t start = (int64_t)fq > prevEnd + 1 ? (int64_t)fq : prevEnd + 1;
    if (start <= (int64_t)e) {  // a run never ends before its quotient in a valid table
      if (EMIT) {
        for (int64_t s = start; s <= (int64_t)e; ++s)
          out[o++] = (K)((fq << f.r) | f.get<0>((uint32_t)(s >> 6), (uint32_t)(s & 63)));
      }
      cnt += (uint32_t)((int64_t)e - start + 1);
    }
    prevEnd = (int64_t)e;
    pos = e + 1;
  }
  if (!EMIT) counts[j] = cnt;
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

static void reset_op_counters_r(hipStream_t s, DeviceState& ds) {
  GQF_LAUNCH(k_reset_counters, dim3(1), dim3(64), 0, s, ds.counters, (int)CTR_SCAN_TILE, (int)CTR_MISC);
}

template <typename K>
static void launch_rwrite(hipStream_t s, DeviceState& ds, const gqf_rsqf& f, const K* ufp, const uint32_t* loc,
                          uint64_t items) {
  const uint32_t groups = (f.numBlocks + RW_BLOCKS - 1) / RW_BLOCKS;
  GQF_LAUNCH_B((double)items * (sizeof(K) + 4) + (double)rsqf_dev_bytes(f.qbits, f.rbits), k_rsqf_write<K>, dim3(groups), dim3(256), 0, s, reinterpret_cast<uint8_t*>(f.blocks), f.numBlocks,
                     f.rbits, ufp, loc, ds.counters);
}

static gqf_status read_ctr(hipStream_t s, DeviceState& ds, const uint32_t* dptr, uint32_t* out) {
  GQF_HIP(hipMemcpyAsync(ds.h_pinned + CTR_MISC + 4, dptr, 4, hipMemcpyDeviceToHost, s));
  GQF_HIP(hipStreamSynchronize(s));
  *out = ds.h_pinned[CTR_MISC + 4];
  return GQF_OK;
}

template <typename K>
static gqf_status rsqf_decode(hipStream_t s, DeviceState& ds, Arena& ar, const gqf_rsqf& f, K** out, uint32_t* m_out) {
  const RsqfDev d = rdev(f);
  uint32_t* counts = ar.take<uint32_t>(f.numBlocks);
  uint32_t* total = ar.take<uint32_t>(1);
  if (ar.used > ar.cap) {
    set_error("internal: decode workspace overflow");
    return GQF_ERR_INVALID;
  }
  const dim3 grid((f.numBlocks + 255) / 256);
  GQF_LAUNCH_B((double)rsqf_meta_bytes(f.numBlocks), (k_rsqf_decode<false, K>), grid, dim3(256), 0, s, d, counts, (K*)nullptr);
  GQF_LAUNCH(k_scan_rows<1>, dim3(1), dim3(1024), 0, s, counts, f.numBlocks, total);
  GQF_TRY(read_ctr(s, ds, total, m_out));
  K* fps = ar.take<K>(*m_out + 1);
  if (ar.used > ar.cap) {
    set_error("internal: decode workspace overflow");
    return GQF_ERR_INVALID;
  }
  GQF_LAUNCH_B((double)rsqf_dev_bytes(f.qbits, f.rbits) + (double)*m_out * sizeof(K), (k_rsqf_decode<true, K>), grid, dim3(256), 0, s, d, counts, fps);
  *out = fps;
  return GQF_OK;
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
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(n, sizeof(K), false) + layout_workspace_bytes(n, sizeof(K))));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters_r(cs.stream, ds);
  K* sorted;
  GQF_TRY(radix_sort<K>(cs.stream, ds, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &sorted, nullptr));
  K* ufp;
  uint32_t* loc;
  GQF_TRY((layout_scan<K, KEEP_ALL>(cs.stream, ds, ar, sorted, nullptr, n, g.q, g.r, 0, g.cap, true, &ufp, &loc)));
  launch_rwrite<K>(cs.stream, ds, f, ufp, loc, n);
  return GQF_OK;
}

static size_t rsqf_decode_ws(const RsqfGeom& g, size_t kb) {
  return aligned_bytes((size_t)g.nb * 4) + 256 + aligned_bytes((g.cap + 1) * kb) + 256;
}

template <typename K>
static gqf_status rsqf_insert_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, rsqf_decode_ws(g, sizeof(K)) + sort_workspace_bytes(n, sizeof(K), false) +
                             aligned_bytes((g.cap + n + 1) * 8) + layout_workspace_bytes(g.cap + n, 8)));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters_r(cs.stream, ds);
  K* A;
  uint32_t m = 0;
  GQF_TRY(rsqf_decode<K>(cs.stream, ds, ar, f, &A, &m));
  K* B;
  GQF_TRY(radix_sort<K>(cs.stream, ds, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &B, nullptr));
  uint64_t* merged = ar.take<uint64_t>((size_t)m + n + 1);
  if (ar.used > ar.cap || (uint64_t)m + n >= GQF_MAX_ITEMS) {
    set_error("table holds %u entries: decode exceeds the workspace (not a canonical table?)", m);
    return GQF_ERR_INVALID;
  }
  launch_merge_tagged<K>(cs.stream, A, m, B, n, merged);
  uint64_t* ufp;
  uint32_t* loc;
  GQF_TRY((layout_scan<uint64_t, KEEP_ALL>(cs.stream, ds, ar, merged, nullptr, m + n, g.q, g.r, 1, g.cap, true, &ufp,
                                           &loc)));
  launch_rwrite<uint64_t>(cs.stream, ds, f, ufp, loc, (uint64_t)m + n);
  return GQF_OK;
}

template <typename K>
static gqf_status rsqf_delete_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, rsqf_decode_ws(g, sizeof(K)) + sort_workspace_bytes(n, sizeof(K), false) +
                             aligned_bytes(g.cap + 1) + layout_workspace_bytes(g.cap, sizeof(K))));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters_r(cs.stream, ds);
  K* A;
  uint32_t m = 0;
  GQF_TRY(rsqf_decode<K>(cs.stream, ds, ar, f, &A, &m));
  K* D;
  GQF_TRY(radix_sort<K>(cs.stream, ds, ar, keys, true, g.bits, nullptr, false, n, 0, g.bits, &D, nullptr));
  uint8_t* keep = ar.take<uint8_t>(m + 1);
  if (ar.used > ar.cap) {
    set_error("internal: delete workspace overflow");
    return GQF_ERR_INVALID;
  }
  if (m) GQF_LAUNCH_B((double)m * (sizeof(K) + 1), k_delete_mark<K>, dim3((m + DM_TILE - 1) / DM_TILE), dim3(DM_THREADS), 0, cs.stream, A, m, D, n, keep);
  K* ufp;
  uint32_t* loc;
  GQF_TRY((layout_scan<K, KEEP_FLAGS>(cs.stream, ds, ar, A, keep, m, g.q, g.r, 0, g.cap, true, &ufp, &loc)));
  launch_rwrite<K>(cs.stream, ds, f, ufp, loc, m);
  return GQF_OK;
}

static void launch_rlookup(hipStream_t s, const gqf_rsqf& f, uint32_t n, const uint32_t* keys, uint32_t* out) {
  const RsqfDev d = rdev(f);
  const dim3 grid((n + 255) / 256);
  if (f.rbits == 8) GQF_LAUNCH_B((double)n * (8 + 128), k_rsqf_lookup<8>, grid, dim3(256), 0, s, d, f.qbits, keys, n, out);
  else GQF_LAUNCH_B((double)n * (8 + 128), k_rsqf_lookup<0>, grid, dim3(256), 0, s, d, f.qbits, keys, n, out);
}

template <typename K>
static gqf_status rsqf_lookup_sorted_t(const gqf_rsqf& f, uint32_t n, const uint32_t* keys, uint32_t* out,
                                       CallScope& cs) {
  DeviceState& ds = *cs.ds;
  const RsqfGeom g = rgeom(f);
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(n, sizeof(K), true)));
  Arena ar(ds.ws, ds.ws_cap);
  K* fps;
  uint32_t* idx;
  // order by the top 16 fingerprint bits (quotient prefix): blocks are then
  // visited in increasing order
  const uint32_t lo = g.bits > 16 ? g.bits - 16 : 0;
  GQF_TRY(radix_sort<K>(cs.stream, ds, ar, keys, true, g.bits, nullptr, true, n, lo, g.bits, &fps, &idx));
  const RsqfDev d = rdev(f);
  const dim3 grid((n + 255) / 256);
  if (f.rbits == 8)
    GQF_LAUNCH_B((double)n * (sizeof(K) + 8 + 128), (k_rsqf_lookup_sorted<8, K>), grid, dim3(256), 0, cs.stream, d, f.qbits, fps, idx, n, out);
  else
    GQF_LAUNCH_B((double)n * (sizeof(K) + 8 + 128), (k_rsqf_lookup_sorted<0, K>), grid, dim3(256), 0, cs.stream, d, f.qbits, fps, idx, n, out);
  return GQF_OK;
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

gqf_status gqf_rsqf_insert(gqf_rsqf cqf, int n, const uint32_t* d_keys, int32_t* d_out, gqf_stream stream, float* ms) {
  GQF_TRY(rsqf_args(cqf, n, d_keys));
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  if (cqf.qbits + cqf.rbits <= 32) GQF_TRY(rsqf_insert_t<uint32_t>(cqf, (uint32_t)n, d_keys, cs));
  else GQF_TRY(rsqf_insert_t<uint64_t>(cqf, (uint32_t)n, d_keys, cs));
  if (d_out && n > 0) launch_rlookup(cs.stream, cqf, (uint32_t)n, d_keys, reinterpret_cast<uint32_t*>(d_out));
  return cs.end(ms, true);
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
    if (!sorted) launch_rlookup(cs.stream, cqf, (uint32_t)n, d_keys, out);
    else if (cqf.qbits + cqf.rbits <= 32) GQF_TRY(rsqf_lookup_sorted_t<uint32_t>(cqf, (uint32_t)n, d_keys, out, cs));
    else GQF_TRY(rsqf_lookup_sorted_t<uint64_t>(cqf, (uint32_t)n, d_keys, out, cs));
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
