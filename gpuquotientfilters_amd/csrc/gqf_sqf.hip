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
#include "gqf_sqf_dev.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

namespace gqf {

// ---------------------------------------------------------------- pipelines
// Packed slot widths ((r+3) % 8 != 0) run every pipeline on a wide copy of
// the table in the workspace: 16-bit elements, the layout of a byte-aligned
// 2-byte table (bpe and bytes below describe it), unpacked from the table
// first where the operation reads it and packed back at the end.
struct SqfGeom {
  uint32_t q, r, bpe, S;
  uint64_t bytes;
  bool packed;
  uint32_t pwords;  // packed widths: words of the caller's table
};
static SqfGeom geom(const gqf_sqf& qf) {
  SqfGeom g;
  g.q = qf.qbits;
  g.r = qf.rbits;
  g.S = sqf_slots(qf.qbits);
  g.packed = sqf_is_packed(qf.rbits);
  g.bpe = g.packed ? 2u : qf.bytesPerElement;
  g.bytes = g.packed ? 2ull * g.S : sqf_bytes(qf.qbits, qf.rbits);
  g.pwords = g.packed ? (uint32_t)sqf_packed_words(qf.qbits, qf.rbits) : 0u;
  return g;
}
static size_t packed_ws_bytes(const SqfGeom& g) { return g.packed ? aligned_bytes(g.bytes) + 256 : 0; }

// The table a pipeline works on: the caller's, or for packed widths a wide
// copy in the workspace (unpacked from the table when `load`).
static gqf_status work_table(hipStream_t s, Arena& ar, const gqf_sqf& qf, const SqfGeom& g, bool load,
                             gqf_sqf* out) {
  *out = qf;
  if (!g.packed) return GQF_OK;
  out->table = ar.take<uint8_t>(g.bytes);
  out->bytesPerElement = 2;
  if (ar.used > ar.cap) {
    set_error("internal: wide-table workspace overflow");
    return GQF_ERR_INVALID;
  }
  if (load)
    GQF_LAUNCH_B((double)g.pwords * 4 + (double)g.bytes, k_sqf_unpack, dim3((g.S + 255) / 256), dim3(256), 0, s,
                 reinterpret_cast<const uint32_t*>(qf.table), g.S, g.r + 3, reinterpret_cast<uint16_t*>(out->table));
  return GQF_OK;
}
// Packs the wide copy back into the caller's table (packed widths only);
// with ctr, not when the call raised the capacity flag.
static void finish_table(hipStream_t s, const gqf_sqf& qf, const gqf_sqf& work, const SqfGeom& g,
                         const uint32_t* ctr) {
  if (!g.packed) return;
  GQF_LAUNCH_B((double)g.bytes + (double)g.pwords * 4, k_sqf_pack, dim3((g.pwords + 255) / 256), dim3(256), 0, s,
               reinterpret_cast<const uint16_t*>(work.table), g.S, g.r + 3, reinterpret_cast<uint32_t*>(qf.table),
               g.pwords, ctr);
}

// Decodes the table into ascending fingerprints (at most S of them); their
// count stays in device memory (*count_dev): later kernels read it, and the
// host sizes their grids for S.
static gqf_status sqf_decode(hipStream_t s, Arena& ar, const gqf_sqf& qf, const SqfGeom& g, uint32_t** fps_out,
                             const uint32_t** count_out) {
  const uint32_t regions = (g.S + DC_SLOTS - 1) / DC_SLOTS;
  uint32_t* counts = ar.take<uint32_t>((size_t)regions * 3);
  uint32_t* count_dev = ar.take<uint32_t>(3);  // column totals; [0] = entries
  uint32_t* fps = ar.take<uint32_t>(g.S);
  if (ar.used > ar.cap) {
    set_error("internal: decode workspace overflow");
    return GQF_ERR_INVALID;
  }
  if (g.bpe == 1) GQF_LAUNCH_B((double)g.bytes, k_sqf_count<1>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  else if (g.bpe == 2) GQF_LAUNCH_B((double)g.bytes, k_sqf_count<2>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  else if (g.bpe == 3) GQF_LAUNCH_B((double)g.bytes, k_sqf_count<3>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  else GQF_LAUNCH_B((double)g.bytes, k_sqf_count<4>, dim3(regions), dim3(256), 0, s, qf.table, g.S, counts);
  // totals: column 0 (entries) is the decoded count
  GQF_TRY(scan_rows<3>(s, ar, counts, regions, count_dev));
  // table read + 4 B per entry written (the count is on the device)
  const double tb = (double)g.bytes;
  if (g.bpe == 1)
    GQF_LAUNCH_BC(tb, count_dev, 4, (k_sqf_decode<1, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  else if (g.bpe == 2)
    GQF_LAUNCH_BC(tb, count_dev, 4, (k_sqf_decode<2, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  else if (g.bpe == 3)
    GQF_LAUNCH_BC(tb, count_dev, 4, (k_sqf_decode<3, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  else
    GQF_LAUNCH_BC(tb, count_dev, 4, (k_sqf_decode<4, uint32_t>), dim3(regions), dim3(256), 0, s, qf.table, g.S, g.r, counts, fps);
  *fps_out = fps;
  *count_out = count_dev;
  return GQF_OK;
}
static size_t sqf_decode_bytes(const SqfGeom& g) {
  const uint32_t regions = (g.S + DC_SLOTS - 1) / DC_SLOTS;
  return aligned_bytes((size_t)regions * 12) + 256 + aligned_bytes((size_t)g.S * 4) + 512 + scan_rows_bytes(regions, 3);
}

// Reads a device counter into host memory (small synchronous readback, used
// where a count sizes the next launches).
static gqf_status read_counter(hipStream_t s, DeviceState& ds, const uint32_t* dptr, uint32_t* out) {
  GQF_HIP(hipMemcpyAsync(ds.h_pinned + CTR_MISC + 4, dptr, 4, hipMemcpyDeviceToHost, s));
  GQF_HIP(hipStreamSynchronize(s));
  *out = ds.h_pinned[CTR_MISC + 4];
  return GQF_OK;
}

static void clear_table(hipStream_t s, const gqf_sqf& qf, const SqfGeom& g) {
  launch_fill<uint32_t>(s, reinterpret_cast<uint32_t*>(qf.table), g.bytes / 4, 0u);
  if (g.bytes % 4) launch_fill<uint8_t>(s, qf.table + (g.bytes & ~3ull), g.bytes % 4, (uint8_t)0);
}

static size_t rebuild_ws_bytes(uint64_t N) {
  const uint64_t tiles = (N + RB_TILE - 1) / RB_TILE;
  return aligned_bytes((tiles + 1) * 4) + aligned_bytes((size_t)RB_MAX_CHUNKS * 16) + 1024 + aligned_bytes(tiles * RB_FIX * 4) +
         aligned_bytes(tiles * 4) + aligned_bytes((N + 1) * 4) + 2048;
}

// Fused rebuild of qf from A (and B when merging) under a keep rule.
// na_dev (may be NULL): A's count on the device; na is then an upper bound
// that sizes the grids, and the sparse-table choice is made on the device.
template <int MODE, bool MERGE>
static gqf_status launch_rebuild(hipStream_t s, uint32_t* counters, Arena& ar, const gqf_sqf& qf, const SqfGeom& g,
                                 const uint32_t* A, uint32_t na, const uint32_t* na_dev, const uint32_t* B,
                                 uint32_t nb, const uint8_t* flags) {
  const uint64_t N = (uint64_t)na + (MERGE ? nb : 0u);
  if (N == 0) {
    GQF_LAUNCH(k_reset_counters, dim3(1), dim3(64), 0, s, counters, (int)CTR_OP, (int)CTR_TOTAL);
    clear_table(s, qf, g);
    return GQF_OK;
  }
  const RbPlan p = rb_plan(N, MERGE ? RB_MERGE_PER : 1u);
  uint32_t* splits = ar.take<uint32_t>(p.nsub + 1);
  uint4* agg = ar.take<uint4>(p.G);
  uint32_t* fix = ar.take<uint32_t>((size_t)p.nsub * RB_FIX);
  uint32_t* fix_cnt = ar.take<uint32_t>(p.nsub);
  uint32_t* fix_more = ar.take<uint32_t>(N + 1);
  if (ar.used > ar.cap) {
    set_error("internal: rebuild workspace overflow");
    return GQF_ERR_INVALID;
  }
  const uint32_t nadd = MERGE ? nb : 0u;
  const double nknown = na_dev ? (double)nadd : (double)N;  // entries the host knows of; the rest: 4 B each on the device
  if (MERGE)
    GQF_LAUNCH((k_rb_splits<uint32_t, false>), dim3(((p.nsub + 1) * RB_SPLIT_LANES + 255) / 256), dim3(256), 0, s, A, na, na_dev, B, nb,
               p.nsub, splits, (const uint32_t*)nullptr);
  GQF_LAUNCH_BC(nknown * 4, na_dev, 4, (k_rb_agg<MODE, MERGE>), dim3(p.G), dim3(RB_THREADS), 0, s, A, na, na_dev, B,
                nb, flags, splits, g.q, g.r, p.chunk, agg, counters);
  GQF_TRY(rb_scan(s, ar, agg, p.G, (uint64_t)g.S, counters, nullptr));
  // sparse tables (N * 8 < S): clear once, then sub-tiles skip windows holding
  // only zeros; otherwise sub-tiles write every slot of their range and the
  // tail is cleared after.  With the count on the device, the kernels decide.
  const int precleared = na_dev ? -1 : (N * 8 < (uint64_t)g.S ? 1 : 0);
  const int pre_mode = na_dev ? 2 : 1, post_mode = na_dev ? 3 : 0;
  if (precleared != 0)
    GQF_LAUNCH(k_sqf_clear_tail, dim3(1024), dim3(256), 0, s, qf.table, g.bpe, g.bytes, counters, pre_mode, na_dev,
               nadd, g.S);
#define GQF_RB_WRITE(BPE_)                                                                                          \
  GQF_LAUNCH_BC(nknown * 4 + (double)g.bytes, na_dev, 4, (k_rb_write<MODE, MERGE, BPE_>), dim3(p.G),               \
                dim3(RB_THREADS), 0, s, qf.table, g.S, g.q, g.r, A, na, na_dev, B, nb, flags, splits, p.chunk, agg, \
                counters, fix, fix_cnt, fix_more, counters + CTR_MISC + 2, (uint32_t)(N + 1), precleared)
  if (g.bpe == 1) GQF_RB_WRITE(1);
  else if (g.bpe == 2) GQF_RB_WRITE(2);
  else if (g.bpe == 3) GQF_RB_WRITE(3);
  else GQF_RB_WRITE(4);
#undef GQF_RB_WRITE
  if (precleared != 1)
    GQF_LAUNCH(k_sqf_clear_tail, dim3(1024), dim3(256), 0, s, qf.table, g.bpe, g.bytes, counters, post_mode, na_dev,
               nadd, g.S);
  GQF_LAUNCH(k_sqf_fixup, dim3(p.nsub + 1), dim3(64), 0, s, qf.table, g.bpe, fix, fix_cnt, fix_more, counters,
             p.nsub, (uint32_t)(N + 1), na_dev, nadd);
  return GQF_OK;
}

static gqf_status check_keys(int n, const void* keys) {
  if (n < 0 || (uint64_t)n >= GQF_MAX_ITEMS || (n > 0 && !keys)) {
    set_error("invalid key array (n=%d)", n);
    return GQF_ERR_INVALID;
  }
  return GQF_OK;
}

// The build's kernels: hash + sort the keys, then the canonical layout
// rebuild, over a workspace arena and the device counters.  A layout that
// does not fit raises the capacity flag in counters[CTR_ERR] before any
// byte of the table is written.
// seg: the sort may take its segmented pass (a call with a stream context).
static size_t sqf_build_ws_bytes(uint32_t n, bool seg) {
  return sort_workspace_bytes(n, 4, false, seg) + rebuild_ws_bytes(n);
}
// ds: the stream's context (null: the sort takes its chunked passes only).
static gqf_status sqf_build_launch(hipStream_t s, Arena& ar, uint32_t* counters, const gqf_sqf& qf, const SqfGeom& g,
                                  uint32_t un, const uint32_t* d_keys, int no_dup, DeviceState* ds) {
  uint32_t* sorted = nullptr;
  GQF_TRY(radix_sort<uint32_t>(s, ar, d_keys, true, (g.q + g.r) | sqf_hash_flag(), nullptr, false, un, 0, g.q + g.r,
                               &sorted, nullptr, ds));
  if (no_dup) return launch_rebuild<KEEP_UNIQUE, false>(s, counters, ar, qf, g, sorted, un, nullptr, nullptr, 0, nullptr);
  return launch_rebuild<KEEP_ALL, false>(s, counters, ar, qf, g, sorted, un, nullptr, nullptr, 0, nullptr);
}

static gqf_status sqf_build_impl(const gqf_sqf& qf, int n, const uint32_t* d_keys, int no_dup, gqf_stream stream,
                                 float* ms) {
  GQF_TRY(sqf_check(qf));
  GQF_TRY(check_keys(n, d_keys));
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  GQF_TRY(ws_reserve(ds, sqf_build_ws_bytes(un, true) + packed_ws_bytes(g)));
  Arena ar(ds.ws, ds.ws_cap);
  gqf_sqf wq;
  GQF_TRY(work_table(cs.stream, ar, qf, g, false, &wq));
  GQF_TRY(sqf_build_launch(cs.stream, ar, ds.counters, wq, g, un, d_keys, no_dup, &ds));
  finish_table(cs.stream, qf, wq, g, ds.counters);
  return cs.end(ms, true);  // a layout that does not fit is detected before any write
}

// ---------------------------------------------------------------- build without allocation
// gqf_sqf_bulk_build_noalloc runs the build's kernels over `scratch`, a
// fixed device array of the library (each GPU holds its own copy from the
// code object's load until the process ends): the counters in its first 256
// bytes, the sort buffers and rebuild states after them.  On the host the
// array's name is only a stand-in for the symbol: the device address comes
// from hipGetSymbolAddress, once per device, and only that address is handed
// to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = 256;
static_assert(CTR_TOTAL * sizeof(uint32_t) <= NOALLOC_COUNTER_BYTES, "counters fit their reserved bytes");
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoallocDevice {
  std::mutex mu;         // one call at a time per device: calls share the array
  void* ptr = nullptr;   // device address of scratch on this device
};
static NoallocDevice g_noalloc[NOALLOC_MAX_DEVICES];

static gqf_status noalloc_device(NoallocDevice** out) {
  int dev = 0;
  GQF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("device %d is beyond the %d devices the no-allocation build supports", dev, NOALLOC_MAX_DEVICES);
    return GQF_ERR_UNSUPPORTED;
  }
  *out = &g_noalloc[dev];
  return GQF_OK;
}
// insert / insertViaMerge / merge: stored entries A merged with sorted B.
// wq: the table worked on (the caller's qf, or its wide copy).
template <int MODE>
static gqf_status sqf_merge_into(CallScope& cs, Arena& ar, const gqf_sqf& qf, const gqf_sqf& wq, const SqfGeom& g,
                                 const uint32_t* A, uint32_t m, const uint32_t* m_dev, const uint32_t* B, uint32_t k,
                                 float* ms) {
  GQF_TRY((launch_rebuild<MODE, true>(cs.stream, cs.ds->counters, ar, wq, g, A, m, m_dev, B, k, nullptr)));
  finish_table(cs.stream, qf, wq, g, cs.ds->counters);
  return cs.end(ms, true);  // a layout that does not fit is detected before any write
}

// Workspace of the localized insert of nb keys.
static size_t sqf_local_bytes(const SqfGeom& g, uint32_t nb) {
  const uint64_t R = (g.S + PI_REGION - 1) / PI_REGION, W = ((uint64_t)g.S + 63) / 64;
  const uint64_t G = ((uint64_t)nb + RB_TILE - 1) / RB_TILE;
  return aligned_bytes(W * 2) + aligned_bytes((R + 1) * 4) + 256 + 5 * aligned_bytes((uint64_t)nb * 4) +
         aligned_bytes(nb) + aligned_bytes(G * 16) + aligned_bytes(CTR_TOTAL * 4) + 256 + aligned_bytes((uint64_t)g.S * 4) +
         aligned_bytes(((G + RBS_PER - 1) / RBS_PER) * 16) + 2048 + scan_rows_bytes((uint32_t)R + 1, 1);
}

// The localized insert (gqf_sqf_dev.hpp, "localized insert") of the sorted
// fingerprints B.  *done: the table was written; otherwise nothing was
// written and the call needs the full rebuild (a region too long, or a key
// past the last empty slot, where the full rebuild decides capacity).
template <int BPE>
static gqf_status sqf_insert_local(CallScope& cs, Arena& ar, const gqf_sqf& qf, const SqfGeom& g, const uint32_t* B,
                                   uint32_t nb, bool* done) {
  hipStream_t s = cs.stream;
  DeviceState& ds = *cs.ds;
  *done = false;
  const uint32_t R = (g.S + PI_REGION - 1) / PI_REGION, W = (g.S + 63) / 64, G = (nb + RB_TILE - 1) / RB_TILE;
  uint16_t* wpre = ar.take<uint16_t>(W);
  uint32_t* rcnt = ar.take<uint32_t>(R + 1);
  uint32_t* tot = ar.take<uint32_t>(4);
  uint32_t* E = ar.take<uint32_t>(nb);
  uint32_t* s0 = ar.take<uint32_t>(nb);
  uint32_t* P = ar.take<uint32_t>(nb);
  uint32_t* J = ar.take<uint32_t>(nb);
  uint32_t* rend = ar.take<uint32_t>(nb);
  uint8_t* head = ar.take<uint8_t>(nb);
  uint4* agg = ar.take<uint4>(G);
  uint32_t* sctr = ar.take<uint32_t>(CTR_TOTAL);
  uint32_t* flags = ar.take<uint32_t>(4);
  uint32_t* region_tmp = ar.take<uint32_t>(g.S);
  if (ar.used > ar.cap) {
    set_error("internal: localized insert workspace overflow");
    return GQF_ERR_INVALID;
  }
  const dim3 kg((nb + 255) / 256);
  GQF_LAUNCH(k_pi_init, dim3(1), dim3(64), 0, s, rcnt + R, flags);
  GQF_LAUNCH_B((double)g.bytes, k_pi_count<BPE>, dim3(R), dim3(256), 0, s, qf.table, g.S, wpre, rcnt);
  GQF_TRY(scan_rows<1>(s, ar, rcnt, R + 1, tot));
  GQF_LAUNCH(k_pi_keys<BPE>, kg, dim3(256), 0, s, qf.table, g.S, g.q, g.r, B, nb, wpre, rcnt, E, s0, flags);
  GQF_LAUNCH(k_pi_jagg, dim3(G), dim3(RB_THREADS), 0, s, E, nb, agg);
  GQF_TRY(rb_scan(s, ar, agg, G, ~0ull, sctr, nullptr));
  GQF_LAUNCH(k_pi_ranks, dim3(G), dim3(RB_THREADS), 0, s, E, nb, agg, rcnt, R, J, flags);
  GQF_LAUNCH(k_pi_pos<BPE>, kg, dim3(256), 0, s, qf.table, g.S, g.q, g.r, B, J, s0, nb, wpre, rcnt, R, P, head);
  GQF_LAUNCH(k_pi_regions, kg, dim3(256), 0, s, s0, P, head, nb, rend, flags);
  GQF_HIP(hipMemcpyAsync(ds.h_pinned, flags, 4, hipMemcpyDeviceToHost, s));
  GQF_HIP(hipStreamSynchronize(s));
  if (ds.h_pinned[0] != 0) return GQF_OK;  // the full rebuild takes over; nothing was written
  GQF_LAUNCH(k_pi_write<BPE>, kg, dim3(256), 0, s, qf.table, g.S, g.r, B, nb, s0, P, head, rend, region_tmp, flags);
  *done = true;
  return GQF_OK;
}

static gqf_status sqf_insert_impl(const gqf_sqf& qf, int n, const uint32_t* d_keys, int mode, gqf_stream stream,
                                  float* ms) {
  GQF_TRY(sqf_check(qf));
  GQF_TRY(check_keys(n, d_keys));
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  const int path = sqf_insert_path();
  // the localized path rewrites byte-aligned tables in place
  const bool local = !g.packed && mode == KEEP_TAGGED_INSERT && un > 0 && path != GQF_INSERT_REBUILD &&
                     (path == GQF_INSERT_LOCAL || (un <= PI_AUTO_MAX && (uint64_t)un * PI_MIN_RATIO <= g.S));
  const size_t full = sqf_decode_bytes(g) + rebuild_ws_bytes((uint64_t)g.S + un) + packed_ws_bytes(g);
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(un, 4, false) + std::max(full, local ? sqf_local_bytes(g, un) : 0)));
  Arena ar(ds.ws, ds.ws_cap);
  uint32_t* B = nullptr;
  GQF_TRY(radix_sort<uint32_t>(cs.stream, ar, d_keys, true, (g.q + g.r) | sqf_hash_flag(), nullptr, false, un, 0, g.q + g.r, &B,
                               nullptr, &ds));
  if (local) {
    const size_t mark = ar.used;
    bool done = false;
    GQF_LAUNCH(k_reset_counters, dim3(1), dim3(64), 0, cs.stream, ds.counters, (int)CTR_OP, (int)CTR_TOTAL);
    if (g.bpe == 1) GQF_TRY(sqf_insert_local<1>(cs, ar, qf, g, B, un, &done));
    else if (g.bpe == 2) GQF_TRY(sqf_insert_local<2>(cs, ar, qf, g, B, un, &done));
    else if (g.bpe == 3) GQF_TRY(sqf_insert_local<3>(cs, ar, qf, g, B, un, &done));
    else GQF_TRY(sqf_insert_local<4>(cs, ar, qf, g, B, un, &done));
    if (done) return cs.end(ms, true);
    ar.used = mark;
  }
  gqf_sqf wq;
  GQF_TRY(work_table(cs.stream, ar, qf, g, true, &wq));
  uint32_t* A;
  const uint32_t* m_dev = nullptr;
  GQF_TRY(sqf_decode(cs.stream, ar, wq, g, &A, &m_dev));
  // at most S stored entries: grids sized for S + n, the count read on the device
  if (mode == KEEP_TAGGED_INSERT)
    return sqf_merge_into<KEEP_TAGGED_INSERT>(cs, ar, qf, wq, g, A, g.S, m_dev, B, un, ms);
  if (mode == KEEP_UNIQUE) return sqf_merge_into<KEEP_UNIQUE>(cs, ar, qf, wq, g, A, g.S, m_dev, B, un, ms);
  return sqf_merge_into<KEEP_ALL>(cs, ar, qf, wq, g, A, g.S, m_dev, B, un, ms);
}

static gqf_status sqf_delete_impl(const gqf_sqf& qf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms) {
  GQF_TRY(sqf_check(qf));
  GQF_TRY(check_keys(n, d_keys));
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  const uint32_t un = (uint32_t)n;
  GQF_TRY(ws_reserve(ds, sort_workspace_bytes(un, 4, false) + sqf_decode_bytes(g) + aligned_bytes(g.S + 1) +
                             rebuild_ws_bytes(g.S) + packed_ws_bytes(g)));
  Arena ar(ds.ws, ds.ws_cap);
  uint32_t* D = nullptr;
  GQF_TRY(radix_sort<uint32_t>(cs.stream, ar, d_keys, true, (g.q + g.r) | sqf_hash_flag(), nullptr, false, un, 0, g.q + g.r, &D,
                               nullptr, &ds));
  gqf_sqf wq;
  GQF_TRY(work_table(cs.stream, ar, qf, g, true, &wq));
  uint32_t* A;
  const uint32_t* m_dev = nullptr;
  GQF_TRY(sqf_decode(cs.stream, ar, wq, g, &A, &m_dev));
  uint8_t* keep = ar.take<uint8_t>((size_t)g.S + 1);
  if (ar.used > ar.cap) {
    set_error("internal: delete workspace overflow");
    return GQF_ERR_INVALID;
  }
  GQF_LAUNCH_BC(0.0, m_dev, 5, (k_delete_mark<uint32_t, false>), dim3((g.S + DM_TILE - 1) / DM_TILE), dim3(DM_THREADS),
                0, cs.stream, A, g.S, m_dev, D, un, keep, (const uint32_t*)nullptr);
  GQF_TRY((launch_rebuild<KEEP_FLAGS, false>(cs.stream, ds.counters, ar, wq, g, A, g.S, m_dev, nullptr, 0, keep)));
  finish_table(cs.stream, qf, wq, g, ds.counters);
  return cs.end(ms, true);  // deleting only shrinks the layout: it always fits
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
  t.bytesPerElement = sqf_is_packed(r) ? 0u : (r + 3) / 8;  // 0: packed (r+3)-bit elements
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

gqf_status gqf_sqf_bulk_build_noalloc(gqf_sqf qf, int method, int n, const uint32_t* d_keys, int no_dup,
                                      gqf_stream stream, float* ms) {
  if (method < GQF_BUILD_SEGMENTED_LAYOUTS || method > GQF_BUILD_SEQUENTIAL_SHIFTS) {
    set_error("unknown build method %d", method);
    return GQF_ERR_INVALID;
  }
  GQF_TRY(sqf_check(qf));
  GQF_TRY(check_keys(n, d_keys));
  if ((uint64_t)n > GQF_NOALLOC_MAX_KEYS) {
    set_error("gqf_sqf_bulk_build_noalloc takes at most %u keys (n=%d); use gqf_sqf_bulk_build",
              (unsigned)GQF_NOALLOC_MAX_KEYS, n);
    return GQF_ERR_INVALID;
  }
  const uint32_t un = (uint32_t)n;
  const SqfGeom g = geom(qf);
  if (NOALLOC_COUNTER_BYTES + sqf_build_ws_bytes(un, false) + packed_ws_bytes(g) > NOALLOC_SCRATCH_BYTES) {
    set_error(g.packed ? "the no-allocation scratch (%zu bytes) cannot hold %u keys plus the wide copy of this "
                         "packed-width table; use gqf_sqf_bulk_build"
                       : "internal: the no-allocation scratch (%zu bytes) is too small for %u keys",
              NOALLOC_SCRATCH_BYTES, un);
    return GQF_ERR_INVALID;
  }
  NoallocDevice* nd = nullptr;
  GQF_TRY(noalloc_device(&nd));
  std::lock_guard<std::mutex> lk(nd->mu);
  if (!nd->ptr) {
    void* p = nullptr;
    GQF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    nd->ptr = p;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(nd->ptr);
  uint32_t* counters = reinterpret_cast<uint32_t*>(base);
  Arena ar(base + NOALLOC_COUNTER_BYTES, NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES);
  // the array keeps the previous call's contents: the counters start at zero
  GQF_LAUNCH(k_reset_counters, dim3(1), dim3(64), 0, s, counters, 0, (int)CTR_TOTAL);
  gqf_sqf wq;
  GQF_TRY(work_table(s, ar, qf, g, false, &wq));
  GQF_TRY(sqf_build_launch(s, ar, counters, wq, g, un, d_keys, no_dup, nullptr));
  finish_table(s, qf, wq, g, counters);
  GQF_HIP(hipGetLastError());
  uint32_t err = 0;
  GQF_HIP(hipMemcpyAsync(&err, counters + CTR_ERR, sizeof(err), hipMemcpyDeviceToHost, s));
  GQF_HIP(hipStreamSynchronize(s));
  if (ms) *ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (err & DEV_ERR_CAPACITY) {
    set_error("capacity exceeded: the entries do not fit the table (table left unchanged)");
    return GQF_ERR_CAPACITY;
  }
  return GQF_OK;
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
  const SqfGeom g = geom(qf);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  const uint32_t un = (uint32_t)n;
  if (un) {
    // 64-slot windows need 16-byte aligned tables (hipMalloc gives 256)
    const bool win = (reinterpret_cast<uintptr_t>(qf.table) & 15u) == 0;
    // The sorted variant (launchSortedLookups) runs the same per-key walk
    // without sorting the keys first: the results are the same by
    // construction, and on MI355X sorting 2^24 keys costs more than the
    // locality it buys (tables up to 147 MB; the RSQF up to 1.6 GB).
    (void)sorted;
    const dim3 grid((un + 255) / 256);
    const double b = (double)un * (8 + 64);
    if (g.packed) GQF_LAUNCH_B(b, k_sqf_lookup_packed, grid, dim3(256), 0, cs.stream, reinterpret_cast<const uint32_t*>(qf.table), g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
    else if (g.bpe == 1 && win) GQF_LAUNCH_B(b, (k_sqf_lookup<1, true>), grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
    else if (g.bpe == 2 && win) GQF_LAUNCH_B(b, (k_sqf_lookup<2, true>), grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
    else if (g.bpe == 1) GQF_LAUNCH_B(b, (k_sqf_lookup<1, false>), grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
    else if (g.bpe == 2) GQF_LAUNCH_B(b, (k_sqf_lookup<2, false>), grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
    else if (g.bpe == 3) GQF_LAUNCH_B(b, (k_sqf_lookup<3, false>), grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
    else GQF_LAUNCH_B(b, (k_sqf_lookup<4, false>), grid, dim3(256), 0, cs.stream, qf.table, g.S, g.q, g.r, d_keys, un, d_out, sqf_hash_flag());
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
  const SqfGeom g = geom(dst);
  CallScope cs;
  GQF_TRY(cs.begin(stream));
  DeviceState& ds = *cs.ds;
  GQF_TRY(ws_reserve(ds, 2 * sqf_decode_bytes(g) + rebuild_ws_bytes(2 * (uint64_t)g.S) + 2 * packed_ws_bytes(g)));
  Arena ar(ds.ws, ds.ws_cap);
  gqf_sqf wd, ws;
  GQF_TRY(work_table(cs.stream, ar, dst, g, true, &wd));
  GQF_TRY(work_table(cs.stream, ar, src, g, true, &ws));
  uint32_t *A, *B;
  const uint32_t *m1_dev = nullptr, *m2_dev = nullptr;
  GQF_TRY(sqf_decode(cs.stream, ar, wd, g, &A, &m1_dev));
  GQF_TRY(sqf_decode(cs.stream, ar, ws, g, &B, &m2_dev));
  uint32_t m2 = 0;  // the second operand's count sizes the merge on the host
  GQF_TRY(read_counter(cs.stream, ds, m2_dev, &m2));
  if (no_dup) return sqf_merge_into<KEEP_UNIQUE>(cs, ar, dst, wd, g, A, g.S, m1_dev, B, m2, ms);
  return sqf_merge_into<KEEP_ALL>(cs, ar, dst, wd, g, A, g.S, m1_dev, B, m2, ms);
}

}  // extern "C"
