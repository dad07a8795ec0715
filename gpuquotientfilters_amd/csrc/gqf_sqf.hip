// gqf_sqf.hip -- standard quotient filter (SQF) engine for gfx950.
//
// Table format is the reference's (quotientFilter.cuh:25-31,
// quotientFilter.cu:151-170): S = (1 << q) * 1.1 slots of (r+3)/8 bytes,
// big-endian, element = remainder<<3 | occupied<<2 | continuation<<1 | shifted.
//
// Every mutating operation produces the canonical layout (runs in quotient
// order, remainders ascending inside a run, entry i at
// loc_i = max(fq_i, loc_{i-1}+1)), which is the end state of the reference's
// builds, supercluster insert/delete and merges.  They all run as
//   [decode table] -> [hash + radix sort keys] -> [merge / mark] ->
//   layout scan (locations) -> table writer (every slot written once).
// Lookups are the reference's run walk (quotientFilter.cu:172-196, 320-357).
#include "gqf_kernels.hpp"

namespace gqf {

constexpr uint32_t SQF_NOT_FOUND = 0xFFFFFFFFu;

// numElements (quotientFilter.cu:174): (1 << q) * 1.1 in double, truncated.
static inline uint32_t sqf_slots(unsigned q) { return (uint32_t)((double)(1u << q) * 1.1); }
// calcNumSlotsGPU (quotientFilter.cu:37-42): bytes.
static inline size_t sqf_bytes(unsigned q, unsigned r) {
  size_t tableBits = (size_t)(1u << q) * (r + 3);
  return (size_t)((double)(tableBits / 8) * 1.1);
}

static gqf_status sqf_check(const gqf_sqf& qf) {
  if (!qf.table) {
    set_error("SQF table is NULL (call gqf_sqf_init)");
    return GQF_ERR_INVALID;
  }
  if (qf.qbits < 1 || qf.rbits < 1 || (qf.rbits + 3) % 8 != 0 || qf.qbits + qf.rbits > 32 || qf.qbits > 28 ||
      qf.bytesPerElement != (qf.rbits + 3) / 8) {
    set_error("unsupported SQF geometry q=%u r=%u bytesPerElement=%u (need (r+3)%%8==0, q+r<=32, q<=28)", qf.qbits,
              qf.rbits, qf.bytesPerElement);
    return GQF_ERR_INVALID;
  }
  return GQF_OK;
}

// ---------------------------------------------------------------- slot access
template <int BPE>
__device__ __forceinline__ uint32_t sqf_get(const uint8_t* __restrict__ t, uint32_t S, uint32_t i) {
  if (i >= S) return 0u;  // past the last slot reads as empty
  if (BPE == 1) return t[i];
  const uint8_t* p = t + (size_t)i * BPE;
  uint32_t e = p[0];
#pragma unroll
  for (int k = 1; k < BPE; ++k) e = (e << 8) | p[k];
  return e;
}

// findRunStartGPU (quotientFilter.cu:172-196) + run scan of hashAndLookUp
// (quotientFilter.cu:333-356): slot of the first matching remainder.
template <int BPE>
__device__ __forceinline__ uint32_t sqf_lookup_one(const uint8_t* __restrict__ t, uint32_t S, uint32_t fq, uint32_t fr) {
  if (!(sqf_get<BPE>(t, S, fq) & 4u)) return SQF_NOT_FOUND;
  uint32_t b = fq;
  while (b > 0 && (sqf_get<BPE>(t, S, b) & 1u)) b--;
  uint32_t s = b;
  while (b != fq) {
    do {
      s++;
    } while ((sqf_get<BPE>(t, S, s) & 2u) && s < S);
    do {
      b++;
    } while (!(sqf_get<BPE>(t, S, b) & 4u) && b < S);
  }
  uint32_t e = sqf_get<BPE>(t, S, s);
  do {
    const uint32_t rem = e >> 3;
    if (rem == fr) return s;
    if (rem > fr) return SQF_NOT_FOUND;
    s++;
    e = sqf_get<BPE>(t, S, s);
  } while (e & 2u);
  return SQF_NOT_FOUND;
}

template <int BPE>
__global__ __launch_bounds__(256) void k_sqf_lookup(const uint8_t* __restrict__ t, uint32_t S, uint32_t q, uint32_t r,
                                                    const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint32_t fp = (uint32_t)fingerprint(keys[i], q + r);
  out[i] = sqf_lookup_one<BPE>(t, S, (fp >> r) & (uint32_t)mask64(q), fp & ((1u << r) - 1u));
}

// Lookups in fingerprint order (sorted path): result scattered to the
// original index.
template <int BPE>
__global__ __launch_bounds__(256) void k_sqf_lookup_sorted(const uint8_t* __restrict__ t, uint32_t S, uint32_t q,
                                                           uint32_t r, const uint32_t* __restrict__ fps,
                                                           const uint32_t* __restrict__ idx, uint32_t n,
                                                           uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint32_t fp = fps[i];
  out[idx[i]] = sqf_lookup_one<BPE>(t, S, (fp >> r) & (uint32_t)mask64(q), fp & ((1u << r) - 1u));
}

// ---------------------------------------------------------------- table writer
// One workgroup per run of WR_SLOTS slots: gathers the entries whose location
// (and the quotients) fall in its range into LDS and writes every byte of the
// range once, zeros included.  Element rules as setMetadata /
// segmentedQFWrite / setOccupiedBits (quotientFilter.cu:700-735, 1430-1443).
constexpr uint32_t WR_SLOTS = 4096;

template <typename T>
__device__ __forceinline__ uint32_t lower_bound_u32(const T* __restrict__ a, uint32_t lo, uint32_t hi, uint64_t v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
template <typename K>
__device__ __forceinline__ uint32_t lower_bound_fq(const K* __restrict__ fp, uint32_t lo, uint32_t hi, uint64_t v,
                                                   uint32_t r, uint64_t mq) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((((uint64_t)fp[mid] >> r) & mq) < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K>
__global__ __launch_bounds__(256) void k_sqf_write(uint8_t* __restrict__ t, uint32_t S, uint64_t tbytes, uint32_t q,
                                                   uint32_t r, uint32_t bpe, const K* __restrict__ ufp,
                                                   const uint32_t* __restrict__ loc, const uint32_t* __restrict__ ctr) {
  if (ctr[CTR_ERR]) return;  // capacity error: the table stays as it was
  __shared__ uint32_t s_el[WR_SLOTS];
  __shared__ uint32_t s_rng[4];
  const uint32_t m = ctr[CTR_COUNT];
  const uint32_t tid = threadIdx.x;
  const uint32_t s0 = blockIdx.x * WR_SLOTS;
  const uint32_t s1 = s0 + WR_SLOTS < S ? s0 + WR_SLOTS : S;
  const uint64_t mq = mask64(q);
  for (uint32_t i = tid; i < WR_SLOTS; i += 256) s_el[i] = 0u;
  if (tid == 0) s_rng[0] = lower_bound_u32(loc, 0u, m, s0);
  if (tid == 64) s_rng[1] = lower_bound_u32(loc, 0u, m, s1);
  if (tid == 128) s_rng[2] = lower_bound_fq(ufp, 0u, m, s0, r, mq);
  if (tid == 192) s_rng[3] = lower_bound_fq(ufp, 0u, m, s1, r, mq);
  __syncthreads();
  for (uint32_t i = s_rng[0] + tid; i < s_rng[1]; i += 256) {
    const uint64_t fp = (uint64_t)ufp[i];
    const uint64_t fq = (fp >> r) & mq;
    const uint32_t lc = loc[i];
    uint32_t e = (uint32_t)(fp & ((1ull << r) - 1ull)) << 3;
    if (i > 0 && ((((uint64_t)ufp[i - 1]) >> r) & mq) == fq) e |= 2u;
    if ((uint64_t)lc != fq) e |= 1u;
    s_el[lc - s0] = e;
  }
  __syncthreads();
  for (uint32_t i = s_rng[2] + tid; i < s_rng[3]; i += 256) {
    const uint64_t fq = ((uint64_t)ufp[i] >> r) & mq;
    if (i == 0 || ((((uint64_t)ufp[i - 1]) >> r) & mq) != fq) s_el[fq - s0] |= 4u;
  }
  __syncthreads();
  // write bytes [s0*bpe, e1) as aligned 32-bit words, big-endian slots
  const uint64_t b0 = (uint64_t)s0 * bpe;
  const bool last = s1 == S;
  const uint64_t e1 = last ? tbytes : (uint64_t)s1 * bpe;
  const uint64_t nb = e1 - b0;
  const uint64_t nwords = nb / 4;
  uint32_t* tw = reinterpret_cast<uint32_t*>(t + b0);
  for (uint64_t w = tid; w < nwords; w += 256) {
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t B = w * 4 + k;
      const uint64_t slot = B / bpe;
      uint32_t byte = 0;
      if (s0 + slot < S) byte = (s_el[slot] >> (8 * (bpe - 1 - (uint32_t)(B % bpe)))) & 0xFFu;
      word |= byte << (8 * k);
    }
    tw[w] = word;
  }
  for (uint64_t B = nwords * 4 + tid; B < nb; B += 256) {
    const uint64_t slot = B / bpe;
    uint32_t byte = 0;
    if (s0 + slot < S) byte = (s_el[slot] >> (8 * (bpe - 1 - (uint32_t)(B % bpe)))) & 0xFFu;
    t[b0 + B] = (uint8_t)byte;
  }
}

// ---------------------------------------------------------------- decode
// Every stored fingerprint in slot (= ascending) order.  The k-th run start of
// the table belongs to the k-th occupied quotient, so a region decodes from
// three exclusive counts (entries, run starts, occupied bits) before it, plus
// the occupied positions just before it that still own runs.  Replaces
// extractQuotients + remove_if (quotientFilter.cu:1561-1600, 1627).
constexpr uint32_t DC_SLOTS = 4096;
constexpr uint32_t DC_PEND = 1024;

template <int BPE>
__global__ __launch_bounds__(256) void k_sqf_count(const uint8_t* __restrict__ t, uint32_t S,
                                                   uint32_t* __restrict__ counts) {
  __shared__ uint64_t s_tmp[4];
  const uint32_t s0 = blockIdx.x * DC_SLOTS + threadIdx.x * 16;
  uint32_t ci = 0, cr = 0, co = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint32_t e = sqf_get<BPE>(t, S, s0 + j);
    ci += (e & 7u) != 0;
    cr += ((e & 7u) != 0) && !(e & 2u);
    co += (e & 4u) != 0;
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
// quotientFilter.cu:1580-1599); fallback for very long pending chains.
template <int BPE>
__device__ uint32_t sqf_walk_quotient(const uint8_t* __restrict__ t, uint32_t S, uint32_t idx) {
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
  __shared__ uint64_t s_tmp[4];
  __shared__ uint32_t s_tmp32[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t s0 = blockIdx.x * DC_SLOTS;
  const uint32_t bI = bases[blockIdx.x * 3 + 0], bR = bases[blockIdx.x * 3 + 1], bO = bases[blockIdx.x * 3 + 2];
  uint32_t e[16];
  uint32_t ci = 0, cr = 0, co = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    e[j] = sqf_get<BPE>(t, S, s0 + tid * 16 + j);
    ci += (e[j] & 7u) != 0;
    cr += ((e[j] & 7u) != 0) && !(e[j] & 2u);
    co += (e[j] & 4u) != 0;
  }
  const uint64_t ex = block256_excl_scan64((uint64_t)ci | ((uint64_t)cr << 21) | ((uint64_t)co << 42), s_tmp, nullptr);
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
    out[bI + xi] = (K)(((uint64_t)fq << r) | (el >> 3));
    xi++;
  }
}

// ---------------------------------------------------------------- pipelines
struct SqfGeom {
  uint32_t q, r, bpe, S;
  uint64_t bytes;
};
static SqfGeom geom(const gqf_sqf& qf) {
  SqfGeom g;
  g.q = qf.qbits;
  g.r = qf.rbits;
  g.bpe = qf.bytesPerElement;
  g.S = sqf_slots(qf.qbits);
  g.bytes = sqf_bytes(qf.qbits, qf.rbits);
  return g;
}

static void reset_op_counters(hipStream_t s, DeviceState& ds) {
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, ds.counters, (int)CTR_SCAN_TILE, (int)CTR_MISC);
}

template <typename K>
static void launch_write(hipStream_t s, DeviceState& ds, const gqf_sqf& qf, const SqfGeom& g, const K* ufp,
                         const uint32_t* loc) {
  const uint32_t regions = (g.S + WR_SLOTS - 1) / WR_SLOTS;
  hipLaunchKernelGGL(k_sqf_write<K>, dim3(regions), dim3(256), 0, s, qf.table, g.S, g.bytes, g.q, g.r, g.bpe, ufp, loc,
                     ds.counters);
}

// decode the table into ascending fingerprints; the count lands in *count_dev
static gqf_status sqf_decode(hipStream_t s, DeviceState& ds, Arena& ar, const gqf_sqf& qf, const SqfGeom& g,
                             uint32_t** fps_out, uint32_t* count_dev) {
  const uint32_t regions = (g.S + DC_SLOTS - 1) / DC_SLOTS;
  uint32_t* counts = ar.take<uint32_t>((size_t)regions * 3);
  uint32_t* fps = ar.take<uint32_t>(g.S);
  if (ar.used > ar.cap) {
    set_error("internal: decode workspace overflow");
    return GQF_ERR_INVALID;
  }
  if (g.bpe == 1) hipLaunchKernelGGL(k_sqf_count<1>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  else if (g.bpe == 2) hipLaunchKernelGGL(k_sqf_count<2>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  else if (g.bpe == 3) hipLaunchKernelGGL(k_sqf_count<3>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  else hipLaunchKernelGGL(k_sqf_count<4>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  // totals: column 0 (entries) is the decoded count
  hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(1024), 0, s, counts, regions, 3u, count_dev);
  if (g.bpe == 1)
    hipLaunchKernelGGL((k_sqf_decode<1, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  else if (g.bpe == 2)
    hipLaunchKernelGGL((k_sqf_decode<2, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  else if (g.bpe == 3)
    hipLaunchKernelGGL((k_sqf_decode<3, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  else
    hipLaunchKernelGGL((k_sqf_decode<4, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  *fps_out = fps;
  return GQF_OK;
}
static size_t sqf_decode_bytes(const SqfGeom& g) {
  return aligned_bytes((size_t)((g.S + DC_SLOTS - 1) / DC_SLOTS) * 12) + aligned_bytes((size_t)g.S * 4) + 512;
}

// Reads a device counter into host memory (small synchronous readback, used
// where a count sizes the next launches).
static gqf_status read_counter(hipStream_t s, DeviceState& ds, const uint32_t* dptr, uint32_t* out) {
  GQF_HIP(hipMemcpyAsync(ds.h_pinned + CTR_MISC + 4, dptr, 4, hipMemcpyDeviceToHost, s));
  GQF_HIP(hipStreamSynchronize(s));
  *out = ds.h_pinned[CTR_MISC + 4];
  return GQF_OK;
}

static gqf_status sqf_build_impl(const gqf_sqf& qf, int n, const uint32_t* d_keys, int no_dup, gqf_stream stream,
                                 float* ms) {
  GQF_TRY(sqf_check(qf));
  if (n < 0 || (uint64_t)n >= GQF_MAX_ITEMS || (n > 0 && !d_keys)) {
    set_error("invalid key array (n=%d)", n);
    return GQF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(api_mutex());
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(un, 4, false) + layout_workspace_bytes(un, 4)));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters(cs.stream, ds);
  uint32_t* sorted = nullptr;
  GQF_TRY(radix_sort<uint32_t>(cs.stream, ds, ar, d_keys, true, g.q + g.r, nullptr, false, un, 0, g.q + g.r, &sorted,
                               nullptr));
  uint32_t *ufp, *loc;
  if (no_dup)
    GQF_TRY((layout_scan<uint32_t, KEEP_UNIQUE>(cs.stream, ds, ar, sorted, nullptr, un, g.q, g.r, 0, g.S, false, &ufp,
                                                &loc)));
  else
    GQF_TRY((layout_scan<uint32_t, KEEP_ALL>(cs.stream, ds, ar, sorted, nullptr, un, g.q, g.r, 0, g.S, false, &ufp,
                                             &loc)));
  launch_write<uint32_t>(cs.stream, ds, qf, g, ufp, loc);
  return cs.end(ms, true);
}

// Shared tail of insert / insertViaMerge / merge: stored entries A (m of
// them), new sorted fingerprints B (k), merged with source tags, then the rule.
static gqf_status sqf_merge_rebuild(hipStream_t s, DeviceState& ds, Arena& ar, const gqf_sqf& qf, const SqfGeom& g,
                                    const uint32_t* A, uint32_t m, const uint32_t* B, uint32_t k, int mode) {
  uint64_t* merged = ar.take<uint64_t>((size_t)m + k + 1);
  if (ar.used > ar.cap) {
    set_error("internal: merge workspace overflow");
    return GQF_ERR_INVALID;
  }
  launch_merge_tagged<uint32_t>(s, A, m, B, k, merged);
  uint64_t* ufp;
  uint32_t* loc;
  const uint32_t total = m + k;
  if (mode == KEEP_TAGGED_INSERT)
    GQF_TRY((layout_scan<uint64_t, KEEP_TAGGED_INSERT>(s, ds, ar, merged, nullptr, total, g.q, g.r, 1, g.S, false, &ufp,
                                                       &loc)));
  else if (mode == KEEP_UNIQUE)
    GQF_TRY((layout_scan<uint64_t, KEEP_UNIQUE>(s, ds, ar, merged, nullptr, total, g.q, g.r, 1, g.S, false, &ufp, &loc)));
  else
    GQF_TRY((layout_scan<uint64_t, KEEP_ALL>(s, ds, ar, merged, nullptr, total, g.q, g.r, 1, g.S, false, &ufp, &loc)));
  launch_write<uint64_t>(s, ds, qf, g, ufp, loc);
  return GQF_OK;
}

static gqf_status sqf_insert_impl(const gqf_sqf& qf, int n, const uint32_t* d_keys, int mode, gqf_stream stream,
                                  float* ms) {
  GQF_TRY(sqf_check(qf));
  if (n < 0 || (uint64_t)n >= GQF_MAX_ITEMS || (n > 0 && !d_keys)) {
    set_error("invalid key array (n=%d)", n);
    return GQF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(api_mutex());
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  GQF_TRY(ws_reserve(ds, sqf_decode_bytes(g) + sort_workspace_bytes(un, 4, false) +
                             aligned_bytes(((size_t)g.S + un + 1) * 8) + layout_workspace_bytes((size_t)g.S + un, 8)));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters(cs.stream, ds);
  uint32_t* A;
  GQF_TRY(sqf_decode(cs.stream, ds, ar, qf, g, &A, ds.counters + CTR_MISC));
  uint32_t m = 0;
  GQF_TRY(read_counter(cs.stream, ds, ds.counters + CTR_MISC, &m));
  uint32_t* B = nullptr;
  GQF_TRY(radix_sort<uint32_t>(cs.stream, ds, ar, d_keys, true, g.q + g.r, nullptr, false, un, 0, g.q + g.r, &B,
                               nullptr));
  GQF_TRY(sqf_merge_rebuild(cs.stream, ds, ar, qf, g, A, m, B, un, mode));
  return cs.end(ms, true);
}

static gqf_status sqf_delete_impl(const gqf_sqf& qf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms) {
  GQF_TRY(sqf_check(qf));
  if (n < 0 || (uint64_t)n >= GQF_MAX_ITEMS || (n > 0 && !d_keys)) {
    set_error("invalid key array (n=%d)", n);
    return GQF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(api_mutex());
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  GQF_TRY(ws_reserve(ds, sqf_decode_bytes(g) + sort_workspace_bytes(un, 4, false) + aligned_bytes(g.S) +
                             layout_workspace_bytes(g.S, 4)));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters(cs.stream, ds);
  uint32_t* A;
  GQF_TRY(sqf_decode(cs.stream, ds, ar, qf, g, &A, ds.counters + CTR_MISC));
  uint32_t m = 0;
  GQF_TRY(read_counter(cs.stream, ds, ds.counters + CTR_MISC, &m));
  uint32_t* D = nullptr;
  GQF_TRY(radix_sort<uint32_t>(cs.stream, ds, ar, d_keys, true, g.q + g.r, nullptr, false, un, 0, g.q + g.r, &D,
                               nullptr));
  uint8_t* keep = ar.take<uint8_t>(m + 1);
  if (ar.used > ar.cap) {
    set_error("internal: delete workspace overflow");
    return GQF_ERR_INVALID;
  }
  if (m) hipLaunchKernelGGL(k_delete_mark<uint32_t>, dim3((m + 255) / 256), dim3(256), 0, cs.stream, A, m, D, un, keep);
  uint32_t *ufp, *loc;
  GQF_TRY((layout_scan<uint32_t, KEEP_FLAGS>(cs.stream, ds, ar, A, keep, m, g.q, g.r, 0, g.S, false, &ufp, &loc)));
  launch_write<uint32_t>(cs.stream, ds, qf, g, ufp, loc);
  return cs.end(ms, true);
}

}  // namespace gqf

using namespace gqf;

extern "C" {

size_t gqf_sqf_table_bytes(unsigned int q, unsigned int r) { return sqf_bytes(q, r); }
uint32_t gqf_sqf_num_slots(unsigned int q) { return sqf_slots(q); }

gqf_status gqf_sqf_init(gqf_sqf* qf, unsigned int q, unsigned int r) {
  if (!qf) return GQF_ERR_INVALID;
  gqf_sqf t;
  t.qbits = q;
  t.rbits = r;
  t.bytesPerElement = (r + 3) / 8;
  t.table = reinterpret_cast<unsigned char*>(1);  // placeholder for the geometry check
  GQF_TRY(sqf_check(t));
  const size_t bytes = sqf_bytes(q, r);
  unsigned char* p = nullptr;
  GQF_HIP(hipMalloc((void**)&p, bytes));
  GQF_HIP(hipMemset(p, 0, bytes));  // the reference leaves it uninitialised (quotientFilter.cu:52-54)
  GQF_HIP(hipDeviceSynchronize());
  t.table = p;
  *qf = t;
  return GQF_OK;
}

gqf_status gqf_sqf_destroy(gqf_sqf* qf) {
  if (!qf) return GQF_ERR_INVALID;
  if (qf->table) GQF_HIP(hipFree(qf->table));
  qf->table = nullptr;
  return GQF_OK;
}

gqf_status gqf_sqf_bulk_build(gqf_sqf qf, int method, int n, const uint32_t* d_keys, int no_dup, gqf_stream stream,
                              float* ms) {
  if (method < GQF_BUILD_SEGMENTED_LAYOUTS || method > GQF_BUILD_SEQUENTIAL_SHIFTS) {
    set_error("unknown build method %d", method);
    return GQF_ERR_INVALID;
  }
  return sqf_build_impl(qf, n, d_keys, no_dup, stream, ms);
}

gqf_status gqf_sqf_insert(gqf_sqf qf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms) {
  return sqf_insert_impl(qf, n, d_keys, KEEP_TAGGED_INSERT, stream, ms);
}

gqf_status gqf_sqf_insert_via_merge(gqf_sqf qf, const uint32_t* d_old, int n_old, const uint32_t* d_new, int n_new,
                                    int no_dup, gqf_stream stream, float* ms) {
  (void)d_old;
  (void)n_old;
  return sqf_insert_impl(qf, n_new, d_new, no_dup ? KEEP_UNIQUE : KEEP_ALL, stream, ms);
}

gqf_status gqf_sqf_delete(gqf_sqf qf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms) {
  return sqf_delete_impl(qf, n, d_keys, stream, ms);
}

gqf_status gqf_sqf_lookup(gqf_sqf qf, int sorted, int n, const uint32_t* d_keys, uint32_t* d_out, gqf_stream stream,
                          float* ms) {
  GQF_TRY(sqf_check(qf));
  if (n < 0 || (uint64_t)n >= GQF_MAX_ITEMS || (n > 0 && (!d_keys || !d_out))) {
    set_error("invalid key/output array (n=%d)", n);
    return GQF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(api_mutex());
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  if (un) {
    if (!sorted) {
      const dim3 grid((un + 255) / 256);
      if (g.bpe == 1) hipLaunchKernelGGL(k_sqf_lookup<1>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out);
      else if (g.bpe == 2) hipLaunchKernelGGL(k_sqf_lookup<2>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out);
      else if (g.bpe == 3) hipLaunchKernelGGL(k_sqf_lookup<3>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out);
      else hipLaunchKernelGGL(k_sqf_lookup<4>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out);
    } else {
      GQF_TRY(ws_reserve(ds, sort_workspace_bytes(un, 4, true)));
      Arena ar(ds.ws, ds.ws_cap);
      uint32_t *fps, *idx;
      // order by the top 16 fingerprint bits: lookups then walk the table in order
      const uint32_t bits = g.q + g.r;
      const uint32_t lo = bits > 16 ? bits - 16 : 0;
      GQF_TRY(radix_sort<uint32_t>(cs.stream, ds, ar, d_keys, true, bits, nullptr, true, un, lo, bits, &fps, &idx));
      const dim3 grid((un + 255) / 256);
      if (g.bpe == 1) hipLaunchKernelGGL(k_sqf_lookup_sorted<1>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, fps, idx, un, d_out);
      else if (g.bpe == 2) hipLaunchKernelGGL(k_sqf_lookup_sorted<2>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, fps, idx, un, d_out);
      else if (g.bpe == 3) hipLaunchKernelGGL(k_sqf_lookup_sorted<3>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, fps, idx, un, d_out);
      else hipLaunchKernelGGL(k_sqf_lookup_sorted<4>, grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, fps, idx, un, d_out);
    }
  }
  return cs.end(ms, false);
}

gqf_status gqf_sqf_merge(gqf_sqf dst, gqf_sqf src, int no_dup, gqf_stream stream, float* ms) {
  GQF_TRY(sqf_check(dst));
  GQF_TRY(sqf_check(src));
  if (dst.qbits != src.qbits || dst.rbits != src.rbits) {
    set_error("filters to merge must have the same q and r (quotientFilter.cu:1749)");
    return GQF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(api_mutex());
  const SqfGeom g = geom(dst);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  GQF_TRY(ws_reserve(ds, 2 * sqf_decode_bytes(g) + aligned_bytes((2 * (size_t)g.S + 1) * 8) +
                             layout_workspace_bytes(2 * (size_t)g.S, 8)));
  Arena ar(ds.ws, ds.ws_cap);
  reset_op_counters(cs.stream, ds);
  uint32_t *A, *B;
  uint32_t m1 = 0, m2 = 0;
  GQF_TRY(sqf_decode(cs.stream, ds, ar, dst, g, &A, ds.counters + CTR_MISC));
  GQF_TRY(sqf_decode(cs.stream, ds, ar, src, g, &B, ds.counters + CTR_MISC + 1));
  GQF_TRY(read_counter(cs.stream, ds, ds.counters + CTR_MISC, &m1));
  GQF_TRY(read_counter(cs.stream, ds, ds.counters + CTR_MISC + 1, &m2));
  GQF_TRY(sqf_merge_rebuild(cs.stream, ds, ar, dst, g, A, m1, B, m2, no_dup ? KEEP_UNIQUE : KEEP_ALL));
  return cs.end(ms, true);
}

}  // extern "C"
