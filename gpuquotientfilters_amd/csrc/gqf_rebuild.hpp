// gqf_rebuild.hpp -- the chunked location scan shared by the SQF and RSQF
// rebuilds: from a sorted sequence of fingerprints (optionally the stable
// merge of two, optionally filtered by a keep rule) to each kept entry's
// canonical location loc_i = max(fq_i, loc_{i-1}+1) and run heads.
//
// Entries are cut into sub-tiles of RB_TILE and chunks of whole sub-tiles; the
// scan runs reduce-then-scan over the chunks (no waits between workgroups):
// a per-chunk aggregate, one exclusive scan of the aggregates, then the
// writer walks each chunk again on top of its prefix.
#pragma once

#include "gqf_kernels.hpp"

namespace gqf {

constexpr int RB_THREADS = 512;
constexpr int RB_ITEMS = 8;
constexpr int RB_WAVES = RB_THREADS / 64;
constexpr int RB_TILE = RB_THREADS * RB_ITEMS;

// Scan state of a block of entries.  Every field is 0 for a block with no
// kept entry and combines by a monotone rule:
//   c    kept count (add);
//   lp   1 + location of the last kept entry when nothing precedes the block
//        (0: none).  The kept entries take distinct locations >= 0, so
//        lp >= c, and the block maps an incoming lp to max(lp + c, lp_block)
//        -- for an incoming 0 as well, where that is lp_block.  The add
//        saturates: a location past the table (a capacity error) never wraps;
//   fq1  1 + quotient of the last kept entry (0: none); kept entries ascend,
//        so it combines by max.  A kept entry starts a run iff its quotient
//        differs from fq1 - 1 before it (0xFFFFFFFF: no entry before it).
struct RS {
  uint32_t c, lp, fq1;
};
__device__ __forceinline__ uint32_t add_sat(uint32_t a, uint32_t b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ RS rs_combine(RS a, RS b) {
  const uint32_t x = add_sat(a.lp, b.c);
  return RS{a.c + b.c, x > b.lp ? x : b.lp, a.fq1 > b.fq1 ? a.fq1 : b.fq1};
}
__device__ __forceinline__ RS rs_none() { return RS{0u, 0u, 0u}; }
// Appends one kept entry of quotient fq: its location is max(fq, last + 1).
__device__ __forceinline__ void rs_push(RS& s, uint32_t fq) {
  const uint32_t a = s.lp + 1u, b = fq + 1u;
  s.lp = a > b ? a : b;
  s.c++;
  s.fq1 = b;
}
// Stored form of a state: {c, lp, fq1, 0}.
__device__ __forceinline__ uint4 rs_pack(RS x) { return make_uint4(x.c, x.lp, x.fq1, 0u); }
__device__ __forceinline__ RS rs_unpack(uint4 w) { return RS{w.x, w.y, w.z}; }

// Exclusive scan across the 64 lanes of a wave (DPP moves, see
// wave_incl_add; lanes without a source read the empty state 0);
// *total = the wave aggregate.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ RS rs_dpp(RS x) {
  return RS{dpp_u32<CTRL, ROW_MASK>(0u, x.c), dpp_u32<CTRL, ROW_MASK>(0u, x.lp), dpp_u32<CTRL, ROW_MASK>(0u, x.fq1)};
}
__device__ __forceinline__ RS rs_wave_excl(RS x, RS* total) {
  RS inc = x;
  inc = rs_combine(rs_dpp<DPP_ROW_SHR | 1>(inc), inc);
  inc = rs_combine(rs_dpp<DPP_ROW_SHR | 2>(inc), inc);
  inc = rs_combine(rs_dpp<DPP_ROW_SHR | 4>(inc), inc);
  inc = rs_combine(rs_dpp<DPP_ROW_SHR | 8>(inc), inc);
  inc = rs_combine(rs_dpp<DPP_ROW_BCAST15, 0xA>(inc), inc);
  inc = rs_combine(rs_dpp<DPP_ROW_BCAST31, 0xC>(inc), inc);
  *total = RS{wave_read(inc.c, 63), wave_read(inc.lp, 63), wave_read(inc.fq1, 63)};
  return rs_dpp<DPP_WAVE_SHR1>(inc);
}

// Waits for every vector memory operation of the thread, global->LDS stage
// loads among them; with a barrier after it, staged LDS data is readable.
__device__ __forceinline__ void rb_stage_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Merge-path split (A entries among the first u*RB_TILE outputs) of every
// sub-tile boundary, RB_SPLIT_LANES lanes per split (group8_search): the
// predicate A[m] <= B[d - 1 - m] holds below the split and fails from it on
// (A first on ties).  na_dev (may be NULL): A's count on the device,
// overriding na.  PK: A is packed (FpArray) with its bounds at hb.
constexpr uint32_t RB_SPLIT_LANES = 8;
template <typename K, bool PK>
__global__ __launch_bounds__(256) void k_rb_splits(const K* __restrict__ A, uint32_t na,
                                                   const uint32_t* __restrict__ na_dev, const K* __restrict__ B,
                                                   uint32_t nb, uint32_t nsub, uint32_t* __restrict__ splits,
                                                   const uint32_t* __restrict__ hb) {
  const uint32_t gt = blockIdx.x * 256 + threadIdx.x;
  const uint32_t u = gt / RB_SPLIT_LANES;
  if (na_dev) na = *na_dev;
  const auto AA = fp_array<K, PK>(A, hb);
  const uint64_t N = (uint64_t)na + nb;
  // every lane stays for the ballots; groups past nsub search an empty range
  const uint32_t d = u <= nsub ? (uint32_t)((uint64_t)u * RB_TILE < N ? (uint64_t)u * RB_TILE : N) : 0u;
  const uint32_t lo = d > nb ? d - nb : 0u, hi = d < na ? d : na;
  const uint32_t a = group8_search(lo, hi, [&](uint32_t m) { return AA[m] <= B[d - 1 - m]; });
  if (gt % RB_SPLIT_LANES == 0 && u <= nsub) splits[u] = a;
}

struct RbPlan {
  uint32_t G, chunk, nsub;  // chunks, entries per chunk (multiple of RB_TILE), sub-tiles
};
constexpr uint32_t RB_MAX_CHUNKS = 65536;
static inline RbPlan rb_plan(uint64_t N, uint32_t min_per = 1) {
  // min_per sub-tiles per workgroup (more where the next sub-tile's input is
  // staged while the current one is processed), more only beyond
  // RB_MAX_CHUNKS chunks
  RbPlan p;
  p.nsub = (uint32_t)((N + RB_TILE - 1) / RB_TILE);
  uint32_t per = (p.nsub + RB_MAX_CHUNKS - 1) / RB_MAX_CHUNKS;
  if (per < min_per) per = min_per;
  p.chunk = per * RB_TILE;
  p.G = (uint32_t)((N + p.chunk - 1) / p.chunk);
  if (p.G == 0) p.G = 1;
  return p;
}

// Block-wide scan of the per-thread states (NW waves): *ex_out = exclusive
// prefix of the calling thread within the block, returns the block
// aggregate.  s_wave: NW + 1 entries (wave aggregates, then their prefixes
// and the total).  Two barriers.
template <int NW>
__device__ __forceinline__ RS rs_block_n(RS a, RS* ex_out, uint4* s_wave) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  RS wt;
  const RS ex = rs_wave_excl(a, &wt);
  if (lane == 0) s_wave[wid] = rs_pack(wt);
  __syncthreads();
  if (wid == 0) {
    const RS x = lane < NW ? rs_unpack(s_wave[lane]) : rs_none();
    RS tot;
    const RS xe = rs_wave_excl(x, &tot);
    if (lane < NW) s_wave[lane] = rs_pack(xe);
    if (lane == 0) s_wave[NW] = rs_pack(tot);
  }
  __syncthreads();
  *ex_out = rs_combine(rs_unpack(s_wave[wid]), ex);
  return rs_unpack(s_wave[NW]);
}
__device__ __forceinline__ RS rs_block(RS a, RS* ex_out, uint4* s_wave) { return rs_block_n<RB_WAVES>(a, ex_out, s_wave); }

// Exclusive scan of the G chunk states (G <= RB_MAX_CHUNKS) in two launches
// of ceil(G / 1024) workgroups, one state per thread with coalesced loads:
//   k_rb_scan_reduce  the aggregate of each workgroup's 1024 states,
//   k_rb_scan_apply   each workgroup combines the aggregates before it (at
//                     most 64, one wave) and rewrites its states as
//                     exclusive prefixes; the last one writes the total.
// The total (also stored in *total when given) fixes the kept count, the
// first slot after the last entry, and a capacity error -- all before any
// table byte is written.
constexpr uint32_t RBS_PER = 1024;
static_assert(RB_MAX_CHUNKS <= 64 * RBS_PER, "the apply kernel combines at most 64 aggregates in one wave");

[[maybe_unused]] static __global__ __launch_bounds__(1024) void k_rb_scan_reduce(const uint4* __restrict__ agg, uint32_t G,
                                                                                 uint4* __restrict__ part) {
  __shared__ uint4 s_wave[16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t i = blockIdx.x * RBS_PER + (uint32_t)tid;
  RS t;
  (void)rs_wave_excl(i < G ? rs_unpack(agg[i]) : rs_none(), &t);
  if (lane == 0) s_wave[wid] = rs_pack(t);
  __syncthreads();
  if (wid == 0) {
    RS tot;
    (void)rs_wave_excl(lane < 16 ? rs_unpack(s_wave[lane]) : rs_none(), &tot);
    if (lane == 0) part[blockIdx.x] = rs_pack(tot);
  }
}

[[maybe_unused]] static __global__ __launch_bounds__(1024) void k_rb_scan_apply(uint4* __restrict__ agg, uint32_t G,
                                                                                const uint4* __restrict__ part,
                                                                                uint64_t S, uint32_t* __restrict__ ctr,
                                                                                uint4* __restrict__ total) {
  __shared__ uint4 s_wave[17];  // wave aggregates, then their prefixes; [16] = carry
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t i = blockIdx.x * RBS_PER + (uint32_t)tid;
  const RS x = i < G ? rs_unpack(agg[i]) : rs_none();
  if (wid == 0) {
    RS carry;
    (void)rs_wave_excl((uint32_t)lane < blockIdx.x ? rs_unpack(part[lane]) : rs_none(), &carry);
    if (lane == 0) s_wave[16] = rs_pack(carry);
  }
  RS wt;
  const RS ex = rs_wave_excl(x, &wt);
  if (lane == 0) s_wave[wid] = rs_pack(wt);
  __syncthreads();
  if (wid == 0) {
    const RS carry = rs_unpack(s_wave[16]);
    RS tot;
    const RS xe = rs_wave_excl(lane < 16 ? rs_unpack(s_wave[lane]) : rs_none(), &tot);
    if (lane < 16) s_wave[lane] = rs_pack(rs_combine(carry, xe));
    if (lane == 0 && blockIdx.x + 1 == gridDim.x) {
      const RS all = rs_combine(carry, tot);
      if (total) *total = rs_pack(all);
      ctr[CTR_COUNT] = all.c;
      ctr[CTR_MISC + 1] = all.lp;                                 // ... unclamped (saturating)
      ctr[CTR_MISC + 3] = (uint32_t)(all.lp < S ? all.lp : S);  // the slot after the last entry
      if (all.lp > S) ctr[CTR_ERR] |= DEV_ERR_CAPACITY;
    }
  }
  __syncthreads();
  if (i < G) agg[i] = rs_pack(rs_combine(rs_unpack(s_wave[wid]), ex));
}

// Host side: the two launches, with their partials taken from the arena.
static inline gqf_status rb_scan(hipStream_t s, Arena& ar, uint4* agg, uint32_t G, uint64_t S, uint32_t* ctr,
                                 uint4* total) {
  const uint32_t P = (G + RBS_PER - 1) / RBS_PER;
  uint4* part = ar.take<uint4>(P ? P : 1);
  if (ar.used > ar.cap) {
    set_error("internal: rebuild workspace overflow");
    return GQF_ERR_INVALID;
  }
  GQF_LAUNCH(k_rb_scan_reduce, dim3(P), dim3(RBS_PER), 0, s, agg, G, part);
  GQF_LAUNCH(k_rb_scan_apply, dim3(P), dim3(RBS_PER), 0, s, agg, G, part, S, ctr, total);
  return GQF_OK;
}

}  // namespace gqf
