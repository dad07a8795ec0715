// gqf_batch.hip -- many standard quotient filters per launch.
//
// A batch holds F SQFs (same r, any q) and, back to back in one device array
// each, their build keys and their lookup keys.  Every stage of the build
// pipeline is one launch over all filters:
//   segmented LSD radix sort   chunks never cross a filter; per 8-bit digit:
//                              upsweep (chunk histograms), row scan (prefix
//                              of every digit over all chunks), per-filter
//                              digit bases (a filter's count of digit d is a
//                              difference of two row prefixes), downsweep
//                              (keys to filter start + digit base + chunk
//                              prefix + rank);
//   rebuild                    sub-tiles of RB_TILE entries never cross a
//                              filter; per-sub-tile location states, one
//                              segmented scan (states reset at each filter's
//                              first sub-tile), the per-filter capacity check,
//                              then the SQF writer on each sub-tile's own
//                              table, the per-filter tail clear and the
//                              occupied-bit fix-ups;
// and all lookups are one launch (each thread finds its filter by binary
// search over the lookup offsets).  The result of every filter is exactly
// what the per-filter calls produce (gqf_sqf_bulk_build with the same keys,
// gqf_sqf_lookup unsorted).
//
// Launch shapes are chosen by an autotuner when the batch is created: keys
// per sort chunk (1, 2, 4 or 8 sub-tiles) and, for lookups, threads per
// workgroup (64..1024) x keys per thread (1, 2, 4), each candidate timed on
// the batch's own inputs with HIP events.  The tuned pipelines are then
// captured as hipGraphs, so a build or a lookup pass is one graph launch.
// No reference counterpart: the reference launches every kernel per filter.
#include "gqf_sqf_dev.hpp"

#include <algorithm>
#include <vector>

namespace gqf {

struct BFilter {
  uint8_t* table;
  uint64_t bytes;   // table bytes
  uint32_t S, q, r, bits;
  uint32_t key0, nkeys;     // build keys [key0, key0 + nkeys) of the batch key array
  uint32_t probe0, nprobe;  // lookup keys
  uint32_t more0;           // this filter's fix-up overflow list (nkeys + 1 entries)
  uint32_t idx0;            // this filter's lookup index: ceil(S / 64) + 1 entries from here
};
struct BChunk {
  uint32_t begin, end;  // batch key indices
  uint32_t f, head;     // filter, index of the filter's first chunk
};
struct BTile {
  uint32_t f, d0, tn, head;  // filter, first entry (relative to the filter), entries, first of its filter
};

// ---------------------------------------------------------------- segmented sort
template <bool HASH>
__global__ __launch_bounds__(256) void k_bs_upsweep(const uint32_t* __restrict__ in, const BChunk* __restrict__ ch,
                                                    const BFilter* __restrict__ flt, uint32_t shift, uint32_t G,
                                                    uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4 * 256];
  const int tid = threadIdx.x, wid = tid >> 6;
  for (int i = tid; i < 4 * 256; i += 256) h[i] = 0;
  __syncthreads();
  const BChunk c = ch[blockIdx.x];
  const uint32_t bits = flt[c.f].bits;
  for (uint32_t t0 = c.begin; t0 < c.end; t0 += SORT_TILE) {
    uint32_t k[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t idx = t0 + (uint32_t)i * 256 + tid;
      k[i] = idx < c.end ? (HASH ? fingerprint32(in[idx], bits) : in[idx]) : 0u;
    }
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t idx = t0 + (uint32_t)i * 256 + tid;
      if (idx < c.end) atomicAdd(&h[wid * 256 + ((k[i] >> shift) & 255u)], 1u);
    }
  }
  __syncthreads();
  hist[(uint64_t)tid * G + blockIdx.x] = h[tid] + h[256 + tid] + h[512 + tid] + h[768 + tid];
}

// dbase[f][d] = keys of filter f with a smaller digit: the filter's count of
// digit d is (row prefix at its next filter's first chunk) - (row prefix at
// its own first chunk), with the row total standing in past the last chunk.
__global__ __launch_bounds__(256) void k_bs_digitbase(const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals,
                                                      const uint32_t* __restrict__ fchunk, uint32_t G,
                                                      uint32_t* __restrict__ dbase) {
  __shared__ uint32_t tmp[4];
  const uint32_t f = blockIdx.x, d = threadIdx.x;
  const uint32_t c0 = fchunk[f], c1 = fchunk[f + 1];
  const uint32_t p0 = c0 < G ? hist[(uint64_t)d * G + c0] : totals[d];
  const uint32_t p1 = c1 < G ? hist[(uint64_t)d * G + c1] : totals[d];
  dbase[(uint64_t)f * 256 + d] = block256_excl_scan(p1 - p0, tmp, nullptr);
}

template <bool HASH>
__global__ __launch_bounds__(256) void k_bs_downsweep(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      const BChunk* __restrict__ ch, const BFilter* __restrict__ flt,
                                                      uint32_t shift, uint32_t G, const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ dbase, uint32_t n_total) {
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_wcnt[4 * 256];
  __shared__ uint32_t s_lstart[256];
  __shared__ uint32_t s_tmp[4];
  __shared__ uint32_t s_keys[SORT_TILE];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const BChunk c = ch[blockIdx.x];
  const BFilter fl = flt[c.f];
  s_run[tid] = fl.key0 + dbase[(uint64_t)c.f * 256 + tid] + hist[(uint64_t)tid * G + blockIdx.x] -
               hist[(uint64_t)tid * G + c.head];
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t t0 = c.begin; t0 < c.end; t0 += SORT_TILE) {
    for (int i = tid; i < 4 * 256; i += 256) s_wcnt[i] = 0;
    __syncthreads();
    const uint32_t base = t0 + (uint32_t)wid * (SORT_ITEMS * 64) + lane;
    uint32_t k[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t idx = base + (uint32_t)i * 64;
      // padding keys (all ones) take the last digit and rank after real keys
      k[i] = idx < c.end ? (HASH ? fingerprint32(in[idx], fl.bits) : in[idx]) : 0xFFFFFFFFu;
    }
    uint32_t rk[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t d = (k[i] >> shift) & 255u;
      const uint64_t m = match_digit(d);
      const uint32_t pre = s_wcnt[wid * 256 + d];
      const uint32_t cnt = (uint32_t)__popcll(m);
      const uint32_t below = (uint32_t)__popcll(m & lt);
      if (below == cnt - 1) s_wcnt[wid * 256 + d] = pre + cnt;
      rk[i] = pre + below;
    }
    __syncthreads();
    uint32_t total;
    {
      const uint32_t d = tid;
      const uint32_t w0 = s_wcnt[d], w1 = s_wcnt[256 + d], w2 = s_wcnt[512 + d], w3 = s_wcnt[768 + d];
      total = w0 + w1 + w2 + w3;
      s_wcnt[d] = 0;
      s_wcnt[256 + d] = w0;
      s_wcnt[512 + d] = w0 + w1;
      s_wcnt[768 + d] = w0 + w1 + w2;
      s_lstart[d] = block256_excl_scan(total, s_tmp, nullptr);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t d = (k[i] >> shift) & 255u;
      s_keys[s_lstart[d] + s_wcnt[wid * 256 + d] + rk[i]] = k[i];
    }
    __syncthreads();
    const uint32_t tile_n = c.end - t0 < (uint32_t)SORT_TILE ? c.end - t0 : (uint32_t)SORT_TILE;
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
      const uint32_t p = tid + i * SORT_THREADS;
      if (p < tile_n) {
        const uint32_t key = s_keys[p];
        const uint32_t d = (key >> shift) & 255u;
        const uint32_t g = s_run[d] + (p - s_lstart[d]);
        if (g < n_total) out[g] = key;  // always true for consistent counts; keeps stores in the buffer
      }
    }
    __syncthreads();
    s_run[tid] += total;
    __syncthreads();
  }
}

// ---------------------------------------------------------------- segmented rebuild
static __global__ void k_bb_reset(uint32_t* __restrict__ a, uint32_t n) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a[i] = 0u;
}

template <int MODE>
__global__ __launch_bounds__(RB_THREADS, 8) void k_bb_agg(const uint32_t* __restrict__ A, const BTile* __restrict__ tl,
                                                          const BFilter* __restrict__ flt, uint4* __restrict__ agg) {
  __shared__ uint4 s_wave[RB_WAVES + 1];
  const BTile t = tl[blockIdx.x];
  const BFilter fl = flt[t.f];
  uint32_t v[RB_ITEMS + 1];
  uint32_t tags, fl_bits;
  rb_load<MODE>(A + fl.key0, nullptr, t.d0, t.tn, v, tags, fl_bits);
  RS a;
  (void)rb_classify<MODE>(v, tags, fl_bits, t.d0, t.tn, fl.r, mask64(fl.q), a);
  RS ex;
  const RS sub = rs_block(a, &ex, s_wave);
  if (threadIdx.x == 0) {
    uint4 w = rs_pack(sub);
    w.w = t.head;
    agg[blockIdx.x] = w;
  }
}

// Segmented states: a flagged element starts a new filter (its prefix is
// none); (a, ha) then (b, hb) = (hb ? b : a then b, ha | hb).
struct RSS {
  RS s;
  uint32_t h;
};
__device__ __forceinline__ RSS rss_combine(RSS a, RSS b) { return RSS{b.h ? b.s : rs_combine(a.s, b.s), a.h | b.h}; }
__device__ __forceinline__ RSS rss_none() { return RSS{rs_none(), 0u}; }
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ RSS rss_dpp(RSS x) {
  RSS y;
  y.s = rs_dpp<CTRL, ROW_MASK>(x.s);
  y.h = dpp_u32<CTRL, ROW_MASK>(0u, x.h);
  return y;
}
__device__ __forceinline__ RSS rss_wave_excl(RSS x, RSS* total) {
  RSS inc = x;
  inc = rss_combine(rss_dpp<DPP_ROW_SHR | 1>(inc), inc);
  inc = rss_combine(rss_dpp<DPP_ROW_SHR | 2>(inc), inc);
  inc = rss_combine(rss_dpp<DPP_ROW_SHR | 4>(inc), inc);
  inc = rss_combine(rss_dpp<DPP_ROW_SHR | 8>(inc), inc);
  inc = rss_combine(rss_dpp<DPP_ROW_BCAST15, 0xA>(inc), inc);
  inc = rss_combine(rss_dpp<DPP_ROW_BCAST31, 0xC>(inc), inc);
  total->s = RS{wave_read(inc.s.c, 63), wave_read(inc.s.lp, 63), wave_read(inc.s.fq1, 63)};
  total->h = wave_read(inc.h, 63);
  return rss_dpp<DPP_WAVE_SHR1>(inc);
}
__device__ __forceinline__ uint4 rss_pack(RSS x) {
  uint4 w = rs_pack(x.s);
  w.w = x.h;
  return w;
}
__device__ __forceinline__ RSS rss_unpack(uint4 w) { return RSS{rs_unpack(w), w.w}; }

// Exclusive segmented scan of the sub-tile states (one workgroup of 1024,
// pieces of 8192 states, 8 per thread); at the last sub-tile of each filter,
// the filter's end location and capacity check.
__global__ __launch_bounds__(1024) void k_bb_scan(uint4* __restrict__ agg, uint32_t G, const BTile* __restrict__ tl,
                                                  const BFilter* __restrict__ flt, uint32_t* __restrict__ ferr,
                                                  uint32_t* __restrict__ fend) {
  constexpr uint32_t SEG = 8, PIECE = 1024 * SEG;
  __shared__ uint4 s_wave[16];
  __shared__ uint4 s_tot;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  RSS carry = rss_none();
  for (uint32_t base = 0; base < G; base += PIECE) {
    const uint32_t i0 = base + (uint32_t)tid * SEG;
    uint4 w[SEG];
#pragma unroll
    for (uint32_t k = 0; k < SEG; ++k) w[k] = i0 + k < G ? agg[i0 + k] : make_uint4(0u, 0u, 0u, 0u);
    RSS seg = rss_none();
#pragma unroll
    for (uint32_t k = 0; k < SEG; ++k) seg = rss_combine(seg, rss_unpack(w[k]));
    RSS wt;
    RSS ex = rss_wave_excl(seg, &wt);
    if (lane == 0) s_wave[wid] = rss_pack(wt);
    __syncthreads();
    if (wid == 0) {
      RSS x = lane < 16 ? rss_unpack(s_wave[lane]) : rss_none();
      RSS tot;
      RSS xe = rss_wave_excl(x, &tot);
      if (lane < 16) s_wave[lane] = rss_pack(rss_combine(carry, xe));
      if (lane == 0) s_tot = rss_pack(rss_combine(carry, tot));
    }
    __syncthreads();
    RSS run = rss_combine(rss_unpack(s_wave[wid]), ex);
#pragma unroll
    for (uint32_t k = 0; k < SEG; ++k) {
      const uint32_t i = i0 + k;
      if (i < G) {
        const RSS e = rss_unpack(w[k]);
        agg[i] = rss_pack(e.h ? rss_none() : run);
        run = rss_combine(run, e);
        if (i + 1 == G || tl[i + 1].head) {
          const uint32_t f = tl[i].f;
          const uint32_t lp = run.s.lp, S = flt[f].S;
          fend[f] = lp < S ? lp : S;  // the slot after the filter's last entry
          if (lp > S) ferr[f] = DEV_ERR_CAPACITY;
        }
      }
    }
    carry = rss_unpack(s_tot);
    __syncthreads();
  }
}

template <int MODE, int BPE>
__global__ __launch_bounds__(RB_THREADS, 4) void k_bb_write(const uint32_t* __restrict__ A, const BTile* __restrict__ tl,
                                                            const BFilter* __restrict__ flt, const uint4* __restrict__ pre,
                                                            const uint32_t* __restrict__ ferr, uint32_t* __restrict__ fix,
                                                            uint32_t* __restrict__ fix_cnt, uint32_t* __restrict__ more,
                                                            uint32_t* __restrict__ more_n) {
  __shared__ __align__(16) uint32_t s_win[rb_win_words<BPE>()];
  __shared__ uint4 s_wave[RB_WAVES + 1];
  __shared__ uint32_t s_nfix;
  const uint32_t u = blockIdx.x;
  const BTile t = tl[u];
  if (ferr[t.f]) return;  // this filter keeps its table (capacity)
  win_clear<BPE>(s_win);  // (the barriers before the first scatter order it)
  const BFilter fl = flt[t.f];
  uint32_t v[RB_ITEMS + 1];
  uint32_t tags, fl_bits;
  rb_load<MODE>(A + fl.key0, nullptr, t.d0, t.tn, v, tags, fl_bits);
  RS a;
  const uint32_t keepbits = rb_classify<MODE>(v, tags, fl_bits, t.d0, t.tn, fl.r, mask64(fl.q), a);
  (void)rb_emit<BPE>(fl.table, fl.q, fl.r, v, keepbits, a, rs_unpack(pre[u]), u, s_wave, s_win, &s_nfix, fix,
                     fix_cnt, more + fl.more0, more_n + t.f, fl.nkeys + 1, 0);
}

// Every filter's slots after its last entry (all of them when it has none).
__global__ __launch_bounds__(256) void k_bb_clear_tail(const BFilter* __restrict__ flt, const uint32_t* __restrict__ ferr,
                                                       const uint32_t* __restrict__ fend, uint32_t bpe) {
  const uint32_t f = blockIdx.y;
  if (ferr[f]) return;
  const BFilter fl = flt[f];
  const uint64_t b0 = (uint64_t)fend[f] * bpe, tb = fl.bytes;
  const uint64_t a0 = (b0 + 15) & ~15ull, a1 = tb & ~15ull;
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x, nth = (uint64_t)gridDim.x * 256;
  uint8_t* t = fl.table;
  if (a0 >= a1) {
    for (uint64_t b = b0 + tid; b < tb; b += nth) t[b] = 0;
    return;
  }
  for (uint64_t b = b0 + tid; b < a0; b += nth) t[b] = 0;
  for (uint64_t b = a1 + tid; b < tb; b += nth) t[b] = 0;
  uint4* w = reinterpret_cast<uint4*>(t + a0);
  const uint64_t nw = (a1 - a0) / 16;
  for (uint64_t i = tid; i < nw; i += nth) w[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Occupied bits of run heads placed outside their sub-tile's range: workgroup
// u < nsub takes sub-tile u's list, workgroup nsub + f filter f's overflow list.
__global__ __launch_bounds__(64) void k_bb_fixup(const BTile* __restrict__ tl, const BFilter* __restrict__ flt,
                                                 const uint32_t* __restrict__ ferr, const uint32_t* __restrict__ fix,
                                                 const uint32_t* __restrict__ fix_cnt, const uint32_t* __restrict__ more,
                                                 const uint32_t* __restrict__ more_n, uint32_t nsub, uint32_t bpe) {
  const uint32_t b = blockIdx.x;
  if (b < nsub) {
    const uint32_t f = tl[b].f;
    if (ferr[f]) return;
    const uint32_t c = fix_cnt[b];
    uint8_t* t = flt[f].table;
    for (uint32_t i = threadIdx.x; i < c; i += 64) sqf_set_occupied(t, bpe, fix[(uint64_t)b * RB_FIX + i]);
  } else {
    const uint32_t f = b - nsub;
    if (ferr[f]) return;
    const BFilter fl = flt[f];
    uint32_t n = more_n[f];
    if (n > fl.nkeys + 1) n = fl.nkeys + 1;
    for (uint32_t i = threadIdx.x; i < n; i += 64) sqf_set_occupied(fl.table, bpe, more[fl.more0 + i]);
  }
}

// ---------------------------------------------------------------- lookup index
// Per 64-slot window of every filter: {occupied bits, run starts} in the
// window, then an exclusive scan per filter, so that entry w holds the
// counts before slot 64 w and the filter's last entry the totals
// (sqf_lookup_indexed).  Thread i handles index entry i; its filter is the
// last f with woff[f] <= i.
template <int BPE, bool WIN>
__global__ __launch_bounds__(256) void k_bi_count(const BFilter* __restrict__ flt, const uint32_t* __restrict__ woff,
                                                  uint32_t F, uint2* __restrict__ idx) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= woff[F]) return;
  uint32_t lo = 0, hi = F;  // woff[lo] <= i < woff[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (woff[mid] <= i) lo = mid;
    else hi = mid;
  }
  const BFilter fl = flt[lo];
  const uint32_t w = i - woff[lo];
  uint32_t o = 0, r = 0;
  const uint32_t s0 = w * 64, s1 = s0 + 64 < fl.S ? s0 + 64 : fl.S;  // the totals entry (s0 >= S) counts nothing
  if (WIN) {
    if (s0 < fl.S) {  // 16-byte loads and bit masks (slots past S read as empty)
      SqfWin<BPE> W;
      sqf_win_load<BPE>(fl.table, fl.S, w, W);
      o = (uint32_t)__popcll(W.occ);
      r = (uint32_t)__popcll((W.occ | W.cont | W.shf) & ~W.cont);
    }
  } else {
    for (uint32_t s = s0; s < s1; ++s) {
      const uint32_t e = sqf_get<BPE>(fl.table, fl.S, s);
      o += (e >> 2) & 1u;
      r += ((e & 7u) != 0 && !(e & 2u)) ? 1u : 0u;
    }
  }
  idx[i] = make_uint2(o, r);
}

// Filter blockIdx.x: exclusive scan of its entries in place, 1024 at a time.
__global__ __launch_bounds__(1024) void k_bi_scan(const uint32_t* __restrict__ woff, uint2* __restrict__ idx) {
  __shared__ uint32_t s_w[2][16];
  const uint32_t f = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t b0 = woff[f], b1 = woff[f + 1];
  uint32_t co = 0, cr = 0;  // carries
  for (uint32_t base = b0; base < b1; base += 1024) {
    const uint32_t i = base + tid;
    const uint2 x = i < b1 ? idx[i] : make_uint2(0u, 0u);
    const uint32_t io = wave_incl_add(x.x), ir = wave_incl_add(x.y);
    if (lane == 63) {
      s_w[0][wid] = io;
      s_w[1][wid] = ir;
    }
    __syncthreads();
    uint32_t po = 0, pr = 0, to = 0, tr = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t a = s_w[0][k], c = s_w[1][k];
      po += (uint32_t)k < wid ? a : 0u;
      pr += (uint32_t)k < wid ? c : 0u;
      to += a;
      tr += c;
    }
    if (i < b1) idx[i] = make_uint2(co + po + io - x.x, cr + pr + ir - x.y);
    co += to;
    cr += tr;
    __syncthreads();
  }
}

// ---------------------------------------------------------------- lookups
// Key i of the batch lookup array: its filter is the last f with
// poff[f] <= i (binary search over F + 1 offsets), then the per-filter
// lookup (unsorted path).  IPT keys per thread, strided by the workgroup.  A
// walk that needs more than BB_WALK_BUDGET windows (a long cluster: hot keys
// of a skewed batch) is finished through the filter's window index; with
// budget 0xFFFFFFFF every walk runs to its end (the per-filter algorithm, no
// index needed).
constexpr uint32_t BB_WALK_BUDGET = 8;
template <int BPE, bool WIN, int IPT>
__global__ __launch_bounds__(1024) void k_bb_lookup(const BFilter* __restrict__ flt, const uint32_t* __restrict__ poff,
                                                    uint32_t F, const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t* __restrict__ out, const uint2* __restrict__ idx,
                                                    uint32_t budget) {
  const uint32_t base = blockIdx.x * blockDim.x * IPT + threadIdx.x;
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + (uint32_t)j * blockDim.x;
    if (i < n) {
      uint32_t lo = 0, hi = F;  // poff[lo] <= i < poff[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (poff[mid] <= i) lo = mid;
        else hi = mid;
      }
      const BFilter fl = flt[lo];
      const uint32_t fp = fingerprint32(keys[i], fl.bits);
      const uint32_t fq = (fp >> fl.r) & (uint32_t)mask64(fl.q), fr = fp & ((1u << fl.r) - 1u);
      uint32_t res = sqf_lookup_fp<BPE, WIN>(fl.table, fl.S, fq, fr, budget);
      if (res == SQF_GIVE_UP) res = sqf_lookup_indexed<BPE>(fl.table, fl.S, idx + fl.idx0, fq, fr);
      out[i] = res;
    }
  }
}

// ---------------------------------------------------------------- host plan
struct SqfBatch {
  int F = 0, no_dup = 0;
  uint32_t r = 0, bpe = 0;
  bool win = false;
  std::vector<gqf_sqf> filters;
  std::vector<uint64_t> koff, poff;
  const uint32_t* d_keys = nullptr;
  const uint32_t* d_probe = nullptr;
  uint32_t* d_out = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;  // created here (graphs cannot be captured on the null stream)
  uint32_t N = 0, P = 0, nsub = 0, npass = 0, gmax = 0;
  // device memory (one allocation)
  char* mem = nullptr;
  BFilter* d_flt = nullptr;
  BTile* d_tl = nullptr;
  BChunk* d_ch = nullptr;
  uint32_t *d_fchunk = nullptr, *d_poff = nullptr, *d_hist = nullptr, *d_tot = nullptr, *d_dbase = nullptr;
  uint32_t *d_k0 = nullptr, *d_k1 = nullptr;
  uint4* d_agg = nullptr;
  uint32_t *d_fix = nullptr, *d_fixcnt = nullptr, *d_more = nullptr, *d_moren = nullptr;
  uint32_t *d_ferr = nullptr, *d_fend = nullptr;
  uint32_t* d_woff = nullptr;  // lookup index offsets (F + 1)
  uint2* d_idx = nullptr;      // lookup index
  uint32_t nidx = 0;
  // plan
  uint32_t G = 0;       // sort chunks of the chosen shape
  int sort_per = 1;     // sub-tiles per sort chunk
  int lk_block = 256, lk_ipt = 1;
  std::vector<float> tune_sort_ms, tune_lookup_ms;
  hipGraphExec_t g_build = nullptr, g_lookup = nullptr;
  int build_nodes = 0, lookup_nodes = 0;
  std::vector<uint32_t> h_err;
};

// Sort chunks for `per` sub-tiles per chunk; returns G and fills first-chunk indices.
static uint32_t batch_chunks(const SqfBatch& b, int per, std::vector<BChunk>& ch, std::vector<uint32_t>& fchunk) {
  ch.clear();
  fchunk.assign((size_t)b.F + 1, 0);
  const uint32_t span = (uint32_t)per * SORT_TILE;
  for (int f = 0; f < b.F; ++f) {
    fchunk[(size_t)f] = (uint32_t)ch.size();
    const uint32_t k0 = (uint32_t)b.koff[(size_t)f], k1 = (uint32_t)b.koff[(size_t)f + 1];
    const uint32_t head = (uint32_t)ch.size();
    for (uint32_t s = k0; s < k1; s += span) ch.push_back({s, std::min(k1, s + span), (uint32_t)f, head});
  }
  fchunk[(size_t)b.F] = (uint32_t)ch.size();
  return (uint32_t)ch.size();
}

static gqf_status batch_upload_chunks(SqfBatch& b, int per) {
  std::vector<BChunk> ch;
  std::vector<uint32_t> fchunk;
  b.G = batch_chunks(b, per, ch, fchunk);
  b.sort_per = per;
  if (b.G) GQF_HIP(hipMemcpyAsync(b.d_ch, ch.data(), ch.size() * sizeof(BChunk), hipMemcpyHostToDevice, b.stream));
  GQF_HIP(hipMemcpyAsync(b.d_fchunk, fchunk.data(), fchunk.size() * 4, hipMemcpyHostToDevice, b.stream));
  GQF_HIP(hipStreamSynchronize(b.stream));
  return GQF_OK;
}

#define GQF_BB_BPE(BPE_VAR, CALL) \
  do {                            \
    switch (BPE_VAR) {            \
      case 1: CALL(1); break;     \
      case 2: CALL(2); break;     \
      case 3: CALL(3); break;     \
      default: CALL(4); break;    \
    }                             \
  } while (0)

// The lookup index of every filter, from its table.
static void batch_index_launch(SqfBatch& b) {
  hipStream_t s = b.stream;
  const dim3 grid((b.nidx + 255) / 256);
#define GQF_BI_COUNT(BPE_, WIN_)                                                                            \
  GQF_LAUNCH_B((double)b.nidx * 64 * (BPE_) + (double)b.nidx * 8, (k_bi_count<BPE_, WIN_>), grid, dim3(256), 0, s, \
               b.d_flt, b.d_woff, (uint32_t)b.F, b.d_idx)
  if (b.win && b.bpe == 1) GQF_BI_COUNT(1, true);
  else if (b.win && b.bpe == 2) GQF_BI_COUNT(2, true);
  else if (b.bpe == 1) GQF_BI_COUNT(1, false);
  else if (b.bpe == 2) GQF_BI_COUNT(2, false);
  else if (b.bpe == 3) GQF_BI_COUNT(3, false);
  else GQF_BI_COUNT(4, false);
#undef GQF_BI_COUNT
  GQF_LAUNCH(k_bi_scan, dim3((unsigned)b.F), dim3(1024), 0, s, b.d_woff, b.d_idx);
}

// The build pipeline: kernel launches only (no host synchronisation), so it
// can be captured.
static void batch_build_launch(SqfBatch& b) {
  hipStream_t s = b.stream;
  const uint32_t F = (uint32_t)b.F;
  GQF_LAUNCH(k_bb_reset, dim3(64), dim3(256), 0, s, b.d_ferr, 3 * F);  // ferr, fend, more_n
  if (b.N == 0) {
    GQF_LAUNCH(k_bb_clear_tail, dim3(64, F), dim3(256), 0, s, b.d_flt, b.d_ferr, b.d_fend, b.bpe);
    return;
  }
  const uint32_t* src = b.d_keys;
  uint32_t* buf[2] = {b.d_k0, b.d_k1};
  for (uint32_t p = 0; p < b.npass; ++p) {
    const uint32_t shift = 8 * p;
    uint32_t* dst = buf[p & 1];
    if (p == 0)
      GQF_LAUNCH_B((double)b.N * 4, k_bs_upsweep<true>, dim3(b.G), dim3(256), 0, s, src, b.d_ch, b.d_flt, shift, b.G,
                   b.d_hist);
    else
      GQF_LAUNCH_B((double)b.N * 4, k_bs_upsweep<false>, dim3(b.G), dim3(256), 0, s, src, b.d_ch, b.d_flt, shift,
                   b.G, b.d_hist);
    GQF_LAUNCH(k_rs_rowscan, dim3(256), dim3(1024), 0, s, b.d_hist, b.G, b.d_tot, (uint32_t*)nullptr, 0u,
               (uint32_t*)nullptr);
    GQF_LAUNCH(k_bs_digitbase, dim3(F), dim3(256), 0, s, b.d_hist, b.d_tot, b.d_fchunk, b.G, b.d_dbase);
    if (p == 0)
      GQF_LAUNCH_B((double)b.N * 8, k_bs_downsweep<true>, dim3(b.G), dim3(256), 0, s, src, dst, b.d_ch, b.d_flt,
                   shift, b.G, b.d_hist, b.d_dbase, b.N);
    else
      GQF_LAUNCH_B((double)b.N * 8, k_bs_downsweep<false>, dim3(b.G), dim3(256), 0, s, src, dst, b.d_ch, b.d_flt,
                   shift, b.G, b.d_hist, b.d_dbase, b.N);
    src = dst;
  }
  const uint32_t* A = src;
  if (b.no_dup)
    GQF_LAUNCH_B((double)b.N * 4, k_bb_agg<KEEP_UNIQUE>, dim3(b.nsub), dim3(RB_THREADS), 0, s, A, b.d_tl, b.d_flt,
                 b.d_agg);
  else
    GQF_LAUNCH_B((double)b.N * 4, k_bb_agg<KEEP_ALL>, dim3(b.nsub), dim3(RB_THREADS), 0, s, A, b.d_tl, b.d_flt,
                 b.d_agg);
  GQF_LAUNCH(k_bb_scan, dim3(1), dim3(1024), 0, s, b.d_agg, b.nsub, b.d_tl, b.d_flt, b.d_ferr, b.d_fend);
  uint64_t tbytes = 0;
  for (const gqf_sqf& q : b.filters) tbytes += gqf_sqf_table_bytes(q.qbits, q.rbits);
#define GQF_BB_WRITE(BPE_)                                                                                        \
  do {                                                                                                            \
    if (b.no_dup)                                                                                                 \
      GQF_LAUNCH_B((double)b.N * 4 + (double)tbytes, (k_bb_write<KEEP_UNIQUE, BPE_>), dim3(b.nsub),               \
                   dim3(RB_THREADS), 0, s, A, b.d_tl, b.d_flt, b.d_agg, b.d_ferr, b.d_fix, b.d_fixcnt, b.d_more,  \
                   b.d_moren);                                                                                    \
    else                                                                                                          \
      GQF_LAUNCH_B((double)b.N * 4 + (double)tbytes, (k_bb_write<KEEP_ALL, BPE_>), dim3(b.nsub),                  \
                   dim3(RB_THREADS), 0, s, A, b.d_tl, b.d_flt, b.d_agg, b.d_ferr, b.d_fix, b.d_fixcnt, b.d_more,  \
                   b.d_moren);                                                                                    \
  } while (0)
  GQF_BB_BPE(b.bpe, GQF_BB_WRITE);
#undef GQF_BB_WRITE
  GQF_LAUNCH(k_bb_clear_tail, dim3(64, F), dim3(256), 0, s, b.d_flt, b.d_ferr, b.d_fend, b.bpe);
  GQF_LAUNCH(k_bb_fixup, dim3(b.nsub + F), dim3(64), 0, s, b.d_tl, b.d_flt, b.d_ferr, b.d_fix, b.d_fixcnt, b.d_more,
             b.d_moren, b.nsub, b.bpe);
}

// The lookups: first the window index of every table as it is now (a
// filter may have changed since the batch build), then one lookup launch.
// walk_only: no index; every walk runs to its end.
static void batch_lookup_launch_as(SqfBatch& b, bool walk_only) {
  if (b.P == 0) return;
  if (!walk_only) batch_index_launch(b);
  const uint32_t budget = walk_only ? 0xFFFFFFFFu : BB_WALK_BUDGET;
  const uint32_t per = (uint32_t)(b.lk_block * b.lk_ipt);
  const dim3 grid((b.P + per - 1) / per), block((unsigned)b.lk_block);
  const double bytes = (double)b.P * (8 + 64);
#define GQF_BB_LK(BPE_, WIN_, IPT_)                                                                                 \
  GQF_LAUNCH_B(bytes, (k_bb_lookup<BPE_, WIN_, IPT_>), grid, block, 0, b.stream, b.d_flt, b.d_poff, (uint32_t)b.F, \
               b.d_probe, b.P, b.d_out, b.d_idx, budget)
#define GQF_BB_LK_IPT(BPE_, WIN_)           \
  do {                                      \
    if (b.lk_ipt == 1) GQF_BB_LK(BPE_, WIN_, 1); \
    else if (b.lk_ipt == 2) GQF_BB_LK(BPE_, WIN_, 2); \
    else GQF_BB_LK(BPE_, WIN_, 4);          \
  } while (0)
  if (b.win && b.bpe == 1) GQF_BB_LK_IPT(1, true);
  else if (b.win && b.bpe == 2) GQF_BB_LK_IPT(2, true);
  else if (b.bpe == 1) GQF_BB_LK_IPT(1, false);
  else if (b.bpe == 2) GQF_BB_LK_IPT(2, false);
  else if (b.bpe == 3) GQF_BB_LK_IPT(3, false);
  else GQF_BB_LK_IPT(4, false);
#undef GQF_BB_LK_IPT
#undef GQF_BB_LK
}
static void batch_lookup_launch(SqfBatch& b) { batch_lookup_launch_as(b, false); }

static gqf_status time_on_stream(SqfBatch& b, void (*fn)(SqfBatch&), int reps, float* best) {
  hipEvent_t e0, e1;
  GQF_HIP(hipEventCreate(&e0));
  GQF_HIP(hipEventCreate(&e1));
  float m = 1e30f;
  for (int i = 0; i < reps; ++i) {
    GQF_HIP(hipEventRecord(e0, b.stream));
    fn(b);
    GQF_HIP(hipEventRecord(e1, b.stream));
    GQF_HIP(hipEventSynchronize(e1));
    GQF_HIP(hipGetLastError());
    float t = 0.f;
    GQF_HIP(hipEventElapsedTime(&t, e0, e1));
    m = std::min(m, t);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *best = m;
  return GQF_OK;
}

static gqf_status capture(SqfBatch& b, void (*fn)(SqfBatch&), hipGraphExec_t* exec, int* nodes) {
  hipGraph_t g = nullptr;
  GQF_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
  fn(b);
  GQF_HIP(hipStreamEndCapture(b.stream, &g));
  size_t n = 0;
  GQF_HIP(hipGraphGetNodes(g, nullptr, &n));
  *nodes = (int)n;
  const hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  GQF_HIP(e);
  return GQF_OK;
}

static void batch_release(SqfBatch& b) {
  if (b.stream) (void)hipStreamSynchronize(b.stream);
  if (b.g_build) (void)hipGraphExecDestroy(b.g_build);
  if (b.g_lookup) (void)hipGraphExecDestroy(b.g_lookup);
  if (b.mem) (void)hipFree(b.mem);
  b.g_build = b.g_lookup = nullptr;
  b.mem = nullptr;
  if (b.own_stream && b.stream) (void)gqf_stream_destroy(reinterpret_cast<gqf_stream>(b.stream));
  b.stream = nullptr;
}

static gqf_status batch_create(SqfBatch* b, const gqf_sqf* filters, int count, int no_dup, const uint32_t* d_keys,
                               const uint64_t* key_offsets, const uint32_t* d_probe, const uint64_t* probe_offsets,
                               uint32_t* d_out, gqf_stream stream) {
  if (!filters || count <= 0) {
    set_error("batch needs at least one filter");
    return GQF_ERR_INVALID;
  }
  b->F = count;
  b->no_dup = no_dup ? 1 : 0;
  b->stream = reinterpret_cast<hipStream_t>(stream);
  if (!b->stream) {
    // a blocking stream: ordered after the caller's work on the null stream
    GQF_HIP(hipStreamCreate(&b->stream));
    b->own_stream = true;
  }
  b->filters.assign(filters, filters + count);
  b->r = filters[0].rbits;
  b->bpe = filters[0].bytesPerElement;
  b->win = b->bpe <= 2;
  for (int f = 0; f < count; ++f) {
    const gqf_sqf& q = filters[f];
    if (q.table && q.rbits >= 1 && (q.rbits + 3) % 8 != 0) {
      set_error("filter %d: packed slot widths ((r+3)%%8 != 0) are not batched; use the per-filter calls", f);
      return GQF_ERR_UNSUPPORTED;
    }
    if (!q.table || q.rbits != b->r || q.bytesPerElement != b->bpe || q.qbits < 1 ||
        q.qbits + q.rbits > 32 || q.qbits > 28) {
      set_error("filter %d: a batch needs initialised SQFs of one r (q + r <= 32)", f);
      return GQF_ERR_INVALID;
    }
    if ((reinterpret_cast<uintptr_t>(q.table) & 15u) != 0) b->win = false;
  }
  auto offsets = [&](const uint64_t* off, const void* dptr, std::vector<uint64_t>& v, uint32_t* total) -> gqf_status {
    v.assign((size_t)count + 1, 0);
    if (!off) {
      *total = 0;
      return GQF_OK;
    }
    for (int f = 0; f <= count; ++f) v[(size_t)f] = off[f];
    for (int f = 0; f < count; ++f)
      if (v[(size_t)f + 1] < v[(size_t)f] || v[(size_t)f + 1] - v[(size_t)f] >= GQF_MAX_ITEMS) {
        set_error("offsets must be non-decreasing with fewer than 2^30 keys per filter");
        return GQF_ERR_INVALID;
      }
    if (v[0] != 0 || v[(size_t)count] >= (1ull << 31) || (v[(size_t)count] > 0 && !dptr)) {
      set_error("offsets must start at 0 and stay below 2^31 keys, with a key array");
      return GQF_ERR_INVALID;
    }
    *total = (uint32_t)v[(size_t)count];
    return GQF_OK;
  };
  GQF_TRY(offsets(key_offsets, d_keys, b->koff, &b->N));
  GQF_TRY(offsets(probe_offsets, d_probe, b->poff, &b->P));
  if (b->P && !d_out) {
    set_error("lookup results array is NULL");
    return GQF_ERR_INVALID;
  }
  b->d_keys = d_keys;
  b->d_probe = d_probe;
  b->d_out = d_out;
  uint32_t maxbits = 1;
  uint64_t nidx = 0;
  std::vector<uint32_t> woff(1, 0u);
  std::vector<BTile> tl;
  std::vector<BFilter> hf((size_t)count);
  for (int f = 0; f < count; ++f) {
    const gqf_sqf& q = filters[f];
    BFilter& x = hf[(size_t)f];
    x.table = q.table;
    x.bytes = gqf_sqf_table_bytes(q.qbits, q.rbits);
    x.S = gqf_sqf_num_slots(q.qbits);
    x.q = q.qbits;
    x.r = q.rbits;
    x.bits = (q.qbits + q.rbits) | sqf_hash_flag();  // width and hash of the fingerprints
    x.key0 = (uint32_t)b->koff[(size_t)f];
    x.nkeys = (uint32_t)(b->koff[(size_t)f + 1] - b->koff[(size_t)f]);
    x.probe0 = (uint32_t)b->poff[(size_t)f];
    x.nprobe = (uint32_t)(b->poff[(size_t)f + 1] - b->poff[(size_t)f]);
    x.more0 = x.key0 + (uint32_t)f;
    x.idx0 = (uint32_t)nidx;
    nidx += (x.S + 63) / 64 + 1;
    woff.push_back((uint32_t)nidx);
    maxbits = std::max(maxbits, q.qbits + q.rbits);
    for (uint32_t d0 = 0; d0 < x.nkeys; d0 += RB_TILE)
      tl.push_back({(uint32_t)f, d0, std::min<uint32_t>(RB_TILE, x.nkeys - d0), d0 == 0 ? 1u : 0u});
  }
  b->nsub = (uint32_t)tl.size();
  b->npass = (maxbits + 7) / 8;
  std::vector<BChunk> ch;
  std::vector<uint32_t> fchunk;
  b->gmax = std::max<uint32_t>(batch_chunks(*b, 1, ch, fchunk), 1);
  // one device allocation for the plan and its workspace
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += aligned_bytes(bytes ? bytes : 1);
    return o;
  };
  const size_t o_flt = take(sizeof(BFilter) * (size_t)count), o_tl = take(sizeof(BTile) * b->nsub),
               o_ch = take(sizeof(BChunk) * b->gmax), o_fchunk = take(4 * ((size_t)count + 1)),
               o_poff = take(4 * ((size_t)count + 1)), o_hist = take((size_t)4 * 256 * b->gmax), o_tot = take(4 * 256),
               o_dbase = take((size_t)4 * 256 * count), o_k0 = take((size_t)4 * b->N), o_k1 = take((size_t)4 * b->N),
               o_agg = take((size_t)16 * b->nsub), o_fix = take((size_t)4 * RB_FIX * b->nsub),
               o_fixcnt = take((size_t)4 * b->nsub), o_more = take((size_t)4 * ((size_t)b->N + count)),
               o_state = take((size_t)4 * 3 * count), o_woff = take(4 * ((size_t)count + 1)),
               o_idx = take((size_t)8 * nidx);
  GQF_HIP(hipMalloc((void**)&b->mem, off));
  char* m = b->mem;
  b->d_flt = reinterpret_cast<BFilter*>(m + o_flt);
  b->d_tl = reinterpret_cast<BTile*>(m + o_tl);
  b->d_ch = reinterpret_cast<BChunk*>(m + o_ch);
  b->d_fchunk = reinterpret_cast<uint32_t*>(m + o_fchunk);
  b->d_poff = reinterpret_cast<uint32_t*>(m + o_poff);
  b->d_hist = reinterpret_cast<uint32_t*>(m + o_hist);
  b->d_tot = reinterpret_cast<uint32_t*>(m + o_tot);
  b->d_dbase = reinterpret_cast<uint32_t*>(m + o_dbase);
  b->d_k0 = reinterpret_cast<uint32_t*>(m + o_k0);
  b->d_k1 = reinterpret_cast<uint32_t*>(m + o_k1);
  b->d_agg = reinterpret_cast<uint4*>(m + o_agg);
  b->d_fix = reinterpret_cast<uint32_t*>(m + o_fix);
  b->d_fixcnt = reinterpret_cast<uint32_t*>(m + o_fixcnt);
  b->d_more = reinterpret_cast<uint32_t*>(m + o_more);
  b->d_ferr = reinterpret_cast<uint32_t*>(m + o_state);
  b->d_fend = b->d_ferr + count;
  b->d_moren = b->d_fend + count;
  b->d_woff = reinterpret_cast<uint32_t*>(m + o_woff);
  b->d_idx = reinterpret_cast<uint2*>(m + o_idx);
  b->nidx = (uint32_t)nidx;
  std::vector<uint32_t> poff32((size_t)count + 1);
  for (int f = 0; f <= count; ++f) poff32[(size_t)f] = (uint32_t)b->poff[(size_t)f];
  GQF_HIP(hipMemcpyAsync(b->d_flt, hf.data(), hf.size() * sizeof(BFilter), hipMemcpyHostToDevice, b->stream));
  if (!tl.empty())
    GQF_HIP(hipMemcpyAsync(b->d_tl, tl.data(), tl.size() * sizeof(BTile), hipMemcpyHostToDevice, b->stream));
  GQF_HIP(hipMemcpyAsync(b->d_poff, poff32.data(), poff32.size() * 4, hipMemcpyHostToDevice, b->stream));
  GQF_HIP(hipMemcpyAsync(b->d_woff, woff.data(), woff.size() * 4, hipMemcpyHostToDevice, b->stream));
  b->h_err.assign((size_t)count, 0);
  // autotune the sort chunk size on the build (the tables are rebuilt from
  // the batch keys each time: the last run leaves them built)
  const int pers[4] = {1, 2, 4, 8};
  float best = 1e30f;
  int best_per = 1;
  for (int per : pers) {
    if (per > 1 && b->gmax / (uint32_t)per < 1024) break;  // too few chunks to fill the chip
    GQF_TRY(batch_upload_chunks(*b, per));
    float t = 0.f;
    GQF_TRY(time_on_stream(*b, batch_build_launch, 2, &t));
    b->tune_sort_ms.push_back(t);
    if (t < best) {
      best = t;
      best_per = per;
    }
  }
  GQF_TRY(batch_upload_chunks(*b, best_per));
  // autotune the lookup launch shape
  if (b->P) {
    const int blocks[5] = {64, 128, 256, 512, 1024}, ipts[3] = {1, 2, 4};
    float lbest = 1e30f;
    int bb = 256, bi = 1;
    for (int blk : blocks)
      for (int ipt : ipts) {
        b->lk_block = blk;
        b->lk_ipt = ipt;
        float t = 0.f;
        GQF_TRY(time_on_stream(*b, batch_lookup_launch, 3, &t));
        b->tune_lookup_ms.push_back(t);
        if (t < lbest) {
          lbest = t;
          bb = blk;
          bi = ipt;
        }
      }
    b->lk_block = bb;
    b->lk_ipt = bi;
  }
  // rebuild with the chosen shape (eager), then capture both pipelines
  GQF_HIP(hipStreamSynchronize(b->stream));
  GQF_TRY(capture(*b, batch_build_launch, &b->g_build, &b->build_nodes));
  if (b->P) GQF_TRY(capture(*b, batch_lookup_launch, &b->g_lookup, &b->lookup_nodes));
  return GQF_OK;
}

}  // namespace gqf

using namespace gqf;

struct gqf_sqf_batch_s {
  SqfBatch b;
};

extern "C" {

gqf_status gqf_sqf_batch_create(gqf_sqf_batch* out, const gqf_sqf* filters, int count, int no_dup,
                                const uint32_t* d_keys, const uint64_t* key_offsets, const uint32_t* d_probe,
                                const uint64_t* probe_offsets, uint32_t* d_out, gqf_stream stream) {
  if (!out) return GQF_ERR_INVALID;
  *out = nullptr;
  gqf_sqf_batch_s* h = new gqf_sqf_batch_s();
  const gqf_status st =
      batch_create(&h->b, filters, count, no_dup, d_keys, key_offsets, d_probe, probe_offsets, d_out, stream);
  if (st != GQF_OK) {
    batch_release(h->b);
    delete h;
    return st;
  }
  *out = h;
  return GQF_OK;
}

gqf_status gqf_sqf_batch_build(gqf_sqf_batch h, int use_graph, int* filter_status, float* ms) {
  if (!h) return GQF_ERR_INVALID;
  SqfBatch& b = h->b;
  CallScope cs;
  GQF_TRY(cs.begin(reinterpret_cast<gqf_stream>(b.stream)));
  if (use_graph) GQF_HIP(hipGraphLaunch(b.g_build, b.stream));
  else batch_build_launch(b);
  GQF_HIP(hipMemcpyAsync(b.h_err.data(), b.d_ferr, (size_t)b.F * 4, hipMemcpyDeviceToHost, b.stream));
  GQF_TRY(cs.end(ms, false));
  bool any = false;
  for (int f = 0; f < b.F; ++f) {
    const int st = b.h_err[(size_t)f] ? GQF_ERR_CAPACITY : GQF_OK;
    if (filter_status) filter_status[f] = st;
    any = any || st != GQF_OK;
  }
  if (any && !filter_status) {
    set_error("capacity exceeded in at least one filter of the batch (those tables left unchanged)");
    return GQF_ERR_CAPACITY;
  }
  return GQF_OK;
}

gqf_status gqf_sqf_batch_lookup(gqf_sqf_batch h, int use_graph, float* ms) {
  if (!h) return GQF_ERR_INVALID;
  SqfBatch& b = h->b;
  CallScope cs;
  GQF_TRY(cs.begin(reinterpret_cast<gqf_stream>(b.stream)));
  if (b.P) {
    if (use_graph) GQF_HIP(hipGraphLaunch(b.g_lookup, b.stream));
    else batch_lookup_launch(b);
  }
  return cs.end(ms, false);
}

gqf_status gqf_sqf_batch_lookup_walk(gqf_sqf_batch h, float* ms) {
  if (!h) return GQF_ERR_INVALID;
  SqfBatch& b = h->b;
  CallScope cs;
  GQF_TRY(cs.begin(reinterpret_cast<gqf_stream>(b.stream)));
  if (b.P) batch_lookup_launch_as(b, true);
  return cs.end(ms, false);
}

gqf_status gqf_sqf_batch_info(gqf_sqf_batch h, gqf_batch_info* info) {
  if (!h || !info) return GQF_ERR_INVALID;
  const SqfBatch& b = h->b;
  gqf_batch_info x{};
  x.filters = b.F;
  x.build_keys = b.N;
  x.lookup_keys = b.P;
  x.sort_passes = (int)b.npass;
  x.sort_chunks = (int)b.G;
  x.sort_subtiles_per_chunk = b.sort_per;
  x.rebuild_subtiles = (int)b.nsub;
  x.lookup_block = b.lk_block;
  x.lookup_keys_per_thread = b.lk_ipt;
  x.build_graph_nodes = b.build_nodes;
  x.lookup_graph_nodes = b.lookup_nodes;
  for (size_t i = 0; i < b.tune_sort_ms.size() && i < 4; ++i) x.tune_build_ms[i] = b.tune_sort_ms[i];
  for (size_t i = 0; i < b.tune_lookup_ms.size() && i < 15; ++i) x.tune_lookup_ms[i] = b.tune_lookup_ms[i];
  *info = x;
  return GQF_OK;
}

gqf_status gqf_sqf_batch_destroy(gqf_sqf_batch h) {
  if (!h) return GQF_OK;
  batch_release(h->b);
  delete h;
  return GQF_OK;
}

}  // extern "C"
