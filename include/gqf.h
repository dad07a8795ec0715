/* gqf.h -- C ABI of the MI355X quotient-filter engine (libgqf.so).
 *
 * Plain C: device pointers, sizes and status codes; no HIP or torch types in
 * any signature (a stream is passed as an opaque pointer; NULL = null stream).
 * Every entry point replaces a reference host function, cited as
 * file:line, where
 *   quotientFilter.cu/.cuh = standardQuotientFilter/GPU_version/
 *   RSQF.cu/.cuh           = rankSelectQuotientFilter/GPU_version/
 *
 * Conventions kept from the reference:
 *   - a filter is one device allocation owned by the caller-visible handle
 *     (freed with gqf_*_destroy, or hipFree on the pointer);
 *   - an all-zero table is an empty filter;
 *   - key arrays are caller-owned device buffers and are never modified;
 *   - lookup results are slot indices, 0xFFFFFFFF (SQF UINT_MAX / RSQF -1) on
 *     a miss;
 *   - every call blocks until its work is done and reports its elapsed
 *     device time in milliseconds through *ms (may be NULL).
 * Differences: status codes instead of asserts/aborts; capacity overflow is
 * reported (GQF_ERR_CAPACITY) and leaves the table unchanged instead of
 * writing out of bounds; key 0xFFFFFFFF is an ordinary key.
 */
#ifndef GQF_H
#define GQF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gqf_status {
  GQF_OK = 0,
  GQF_ERR_INVALID = 1,   /* bad argument: q/r out of range, NULL pointer, n < 0 */
  GQF_ERR_CAPACITY = 2,  /* the result does not fit the table; table left unchanged */
  GQF_ERR_HIP = 3,       /* HIP runtime error (message in gqf_last_error) */
  GQF_ERR_NOMEM = 4,     /* device allocation failed */
  GQF_ERR_UNSUPPORTED = 5
} gqf_status;

#define GQF_NOT_FOUND 0xFFFFFFFFu

/* Opaque stream handle (a hipStream_t); NULL selects the null stream. */
typedef void* gqf_stream;

/* Last error message of the calling thread (empty string if none). */
const char* gqf_last_error(void);
const char* gqf_status_string(gqf_status s);
/* Library version string, e.g. "gqf 0.1 gfx950". */
const char* gqf_version(void);

/* ====================================================================== SQF
 * Standard quotient filter.  Same layout as the reference struct
 * quotient_filter (quotientFilter.cuh:25-31) and the same table format:
 * floor(2^q * 1.1) slots of (r+3)/8 bytes, big-endian, element =
 * remainder<<3 | occupied<<2 | continuation<<1 | shifted.
 * Supported: q >= 1, q <= 28, q + r <= 32, and either (r + 3) % 8 == 0
 * (r = 5 or 13: the reference GPU format) or, beyond the reference GPU code
 * (README.md:32-33), any r with r + 3 <= 16: the same S slots and elements,
 * (r+3) bits each, packed MSB-first into 32-bit words in the CPU_version's
 * order (getElement, quotientFilter.cpp:114-130), plus one trailing word;
 * such a filter has bytesPerElement == 0.  Every SQF call accepts both
 * layouts (batches: byte-aligned ones only), with the same results for the
 * same q, r and keys. */
typedef struct gqf_sqf {
  unsigned int qbits;
  unsigned int rbits;
  unsigned int bytesPerElement;
  unsigned char* table;
} gqf_sqf;

/* Hash of the SQF calls made from the calling thread: AP (Normal_APHash,
 * quotientFilter.cu:134-149; the reference GPU code's choice, default) or FNV
 * (FNVhashGPU, quotientFilter.cu:117-132, present but unused there; the
 * reference CPU_version's hash, so FNV tables equal the CPU_version's).  A
 * filter must be built, updated and queried with one hash.  Batches take the
 * hash current at gqf_sqf_batch_create.  The RSQF always uses AP. */
enum gqf_hash { GQF_HASH_AP = 0, GQF_HASH_FNV = 1 };
gqf_status gqf_sqf_set_hash(int kind);
int gqf_sqf_get_hash(void);

/* How gqf_sqf_insert (insert, quotientFilter.cu:1184) of the calling thread
 * updates the table.  No reference counterpart: the results are the same
 * either way; only the work differs.
 *   AUTO     localized for batches of at most 65536 keys (and S/256), else
 *            the rebuild;
 *   REBUILD  decode the table, merge the sorted batch, rebuild every slot;
 *   LOCAL    rewrite only the regions the batch changes, between the last
 *            empty-or-unshifted slot before each key and the last slot its
 *            keys take; a region longer than 2^16 slots, or a batch that
 *            does not fit, falls back to the rebuild. */
enum gqf_insert_path { GQF_INSERT_AUTO = 0, GQF_INSERT_REBUILD = 1, GQF_INSERT_LOCAL = 2 };
gqf_status gqf_sqf_set_insert_path(int path);
int gqf_sqf_get_insert_path(void);
/* The same choice for this thread's RSQF inserts (no reference counterpart;
 * the results are equal, only the work differs):
 *   AUTO, LOCAL  rewrite the table in place, window by window (4096 slots:
 *            each window's stored entries decoded from its run ends, merged
 *            with its batch entries, laid out from its incoming location);
 *            a window of more than 8192 merged entries, or runs crossing a
 *            whole window, sends the call to the rebuild before any write;
 *            tables of more than 2^16 windows (q > 27) always rebuild;
 *   REBUILD  decode the table, merge the sorted batch, rebuild every slot. */
gqf_status gqf_rsqf_set_insert_path(int path);
int gqf_rsqf_get_insert_path(void);

enum gqf_sqf_build_method {
  GQF_BUILD_SEGMENTED_LAYOUTS = 0,  /* bulkBuildSegmentedLayouts (quotientFilter.cu:1446) */
  GQF_BUILD_PARALLEL_MERGING = 1,   /* bulkBuildParallelMerging  (quotientFilter.cu:737)  */
  GQF_BUILD_SEQUENTIAL_SHIFTS = 2   /* bulkBuildSequentialShifts (quotientFilter.cu:930)
                                       == README "bulkBuildEarlyExit" */
};

/* calcNumSlotsGPU (quotientFilter.cu:37-42): table size in BYTES. */
size_t gqf_sqf_table_bytes(unsigned int q, unsigned int r);
/* Number of slots: (1 << q) * 1.1 (quotientFilter.cu:174). */
uint32_t gqf_sqf_num_slots(unsigned int q);

/* initFilterGPU (quotientFilter.cu:44-55).  Allocates and zeroes the table. */
gqf_status gqf_sqf_init(gqf_sqf* qf, unsigned int q, unsigned int r);
/* hipFree of qf->table; sets it to NULL. */
gqf_status gqf_sqf_destroy(gqf_sqf* qf);

/* bulkBuild{SegmentedLayouts,ParallelMerging,SequentialShifts}
 * (quotientFilter.cu:1446-1559, 737-878, 930-1058).  Replaces the table
 * contents with the canonical layout of the n keys' fingerprints
 * (duplicates removed when no_dup != 0). */
gqf_status gqf_sqf_bulk_build(gqf_sqf qf, int method, int n, const uint32_t* d_keys, int no_dup,
                              gqf_stream stream, float* ms);

/* No reference counterpart: gqf_sqf_bulk_build for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results (the same
 * table, GQF_ERR_CAPACITY with the table unchanged); it calls no hipMalloc,
 * hipFree or other allocating routine and needs no stream context: its
 * sort buffers, rebuild states and counters live in a fixed 32 MiB device
 * array of the library, one per GPU, held from the library's load until the
 * process ends.  n <= GQF_NOALLOC_MAX_KEYS (larger n: GQF_ERR_INVALID, no
 * device work).  Every call on a GPU shares that GPU's array: the library
 * serialises calls per device (each blocks until its kernels finish), so
 * they cannot overlap within a process.  *ms is the call's wall time. */
#define GQF_NOALLOC_MAX_KEYS (1u << 21)
gqf_status gqf_sqf_bulk_build_noalloc(gqf_sqf qf, int method, int n, const uint32_t* d_keys, int no_dup,
                                      gqf_stream stream, float* ms);

/* insert (quotientFilter.cu:533-634).  Deduplicating insert: a fingerprint
 * already stored is not added again. */
gqf_status gqf_sqf_insert(gqf_sqf qf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms);

/* insertViaMerge (quotientFilter.cu:1602-1742).  Extract the stored
 * fingerprints, merge with the new keys', rebuild.  d_old/n_old are accepted
 * for signature compatibility and not read (the reference does not read them
 * either). */
gqf_status gqf_sqf_insert_via_merge(gqf_sqf qf, const uint32_t* d_old, int n_old,
                                    const uint32_t* d_new, int n_new, int no_dup,
                                    gqf_stream stream, float* ms);

/* launchUnsortedLookups (quotientFilter.cu:359-381) when sorted == 0,
 * launchSortedLookups (quotientFilter.cu:272-318) when sorted != 0.  Both
 * write, for each key, the slot of the first matching remainder of its run,
 * or GQF_NOT_FOUND.  Both run the same per-key walk: the reference sorts the
 * keys first for locality, which on MI355X costs more than it saves (the
 * results do not depend on the order). */
gqf_status gqf_sqf_lookup(gqf_sqf qf, int sorted, int n, const uint32_t* d_keys, uint32_t* d_out,
                          gqf_stream stream, float* ms);

/* superclusterDeletes (quotientFilter.cu:1202-1297).  Removes one stored
 * copy per key whose fingerprint is present. */
gqf_status gqf_sqf_delete(gqf_sqf qf, int n, const uint32_t* d_keys, gqf_stream stream, float* ms);

/* mergeTwoFilters (quotientFilter.cu:1744-1891).  dst := merge of dst and
 * src (same q and r); src unchanged. */
gqf_status gqf_sqf_merge(gqf_sqf dst, gqf_sqf src, int no_dup, gqf_stream stream, float* ms);

/* ===================================================================== RSQF
 * Rank-and-select quotient filter.  The first three fields match the
 * reference struct countingQuotientFilterGPU (RSQF.cuh:32-37); rbits is
 * appended because the reference fixes RBITS at compile time.  The device
 * layout is the project's (see gqf_rsqf_table_bytes); gqf_rsqf_dump/load
 * convert to and from the reference packed block layout
 * {u8 offset; u64 occupieds; u64 runEnds; u64 remainders[r]} (RSQF.cuh:24-30).
 * Supported: 1 <= r <= 16, 6 <= q <= 30. */
typedef struct gqf_rsqf {
  unsigned int qbits;
  unsigned int numBlocks;
  void* blocks;
  unsigned int rbits;
} gqf_rsqf;

/* calcNumBlocksGPU (RSQF.cu:27-30): ceil(2^q / 64) + 1. */
uint32_t gqf_rsqf_num_blocks(unsigned int q);
/* Device bytes of the project layout. */
size_t gqf_rsqf_table_bytes(unsigned int q, unsigned int r);
/* Bytes of the reference packed layout (what dump/load exchange). */
size_t gqf_rsqf_ref_bytes(unsigned int q, unsigned int r);

/* initCQFGPU (RSQF.cu:32-40): allocate and zero. */
gqf_status gqf_rsqf_init(gqf_rsqf* cqf, unsigned int q, unsigned int r);
gqf_status gqf_rsqf_destroy(gqf_rsqf* cqf);

/* Bulk build (no reference counterpart; README.md:74 names it as missing).
 * Replaces the contents with the canonical layout of the keys' fingerprints
 * (duplicates kept, as the reference insert keeps them). */
gqf_status gqf_rsqf_bulk_build(gqf_rsqf cqf, int n, const uint32_t* d_keys, gqf_stream stream,
                               float* ms);

/* insertGPU (RSQF.cu:882-988).  End state: canonical layout of the stored
 * multiset plus the new fingerprints.  d_out (may be NULL) receives, per key,
 * its slot in the resulting table (what a lookup of that key returns). */
gqf_status gqf_rsqf_insert(gqf_rsqf cqf, int n, const uint32_t* d_keys, int32_t* d_out,
                           gqf_stream stream, float* ms);

/* insertGPU with the reference's per-item failure report (RSQF.cu:745,753,
 * 816,860 write -1 for an item that cannot be placed; testGPURSQF.cu:71-80
 * counts them).  A batch that fits is inserted as by gqf_rsqf_insert.  One
 * that does not is inserted in part: with e = 1 + the last location of the
 * whole merge and S = 64 numBlocks slots, the keys whose fingerprints are
 * among the e - S largest of the batch -- every copy of the smallest such
 * fingerprint included -- are left out, the rest is inserted (it fits:
 * every entry before them moves back by e - S, only stored entries follow
 * them).  d_out (may be NULL): slots as gqf_rsqf_insert, -1 for the keys
 * left out; *n_failed (may be NULL): their count.  GQF_OK in both cases.
 * Which items fail in the reference depends on thread timing; this rule is
 * deterministic. */
gqf_status gqf_rsqf_insert_report(gqf_rsqf cqf, int n, const uint32_t* d_keys, int32_t* d_out, uint32_t* n_failed,
                                  gqf_stream stream, float* ms);

/* launchUnsortedLookups (RSQF.cu:540-563) when sorted == 0, launchLookups
 * (RSQF.cu:411-456) when sorted != 0: the LAST matching slot of the run,
 * or -1.  Both run the same per-key walk (no sort first: on MI355X sorting
 * costs more than its locality saves, measured up to 1.6 GB tables).  The walk is the reference's and reads the 8-bit block offsets,
 * which the format stores mod 256 (RSQF.cu offset++ on a uchar): in the
 * blocks behind a run that pushes an offset past 255 a lookup can land on
 * the wrong run and return -1 for a stored key, exactly as the reference
 * does. */
gqf_status gqf_rsqf_lookup(gqf_rsqf cqf, int sorted, int n, const uint32_t* d_keys,
                           int32_t* d_out, gqf_stream stream, float* ms);

/* Counting query (no reference counterpart; README.md:74 names counters as
 * missing): per key, the number of stored copies of its fingerprint -- the
 * RSQF here keeps a multiset, so the copies are the slots of the key's run
 * holding its remainder.  The run is found from run ranks (the k-th run end
 * closes the run of the k-th occupied quotient; per-block prefix counts of
 * both bit vectors are built first), not from the 8-bit block offsets, so
 * counts are exact for runs of any length, also behind runs that push an
 * offset past 255. */
gqf_status gqf_rsqf_count(gqf_rsqf cqf, int n, const uint32_t* d_keys, uint32_t* d_counts, gqf_stream stream,
                          float* ms);

/* Delete (no reference counterpart): removes one stored copy per key whose
 * fingerprint is present; result is the canonical layout of what remains. */
gqf_status gqf_rsqf_delete(gqf_rsqf cqf, int n, const uint32_t* d_keys, gqf_stream stream,
                           float* ms);

/* Device workspace (bytes) a call of `op` (gqf_op) over n keys reserves on
 * its stream's context (lookups: none; `sorted` is kept for the signature).
 * No reference counterpart (the reference allocates per call). */
size_t gqf_rsqf_workspace_bytes(gqf_rsqf cqf, int op, int n, int sorted);

/* Copy the table to host memory in the reference packed layout
 * (gqf_rsqf_ref_bytes bytes), and back. */
gqf_status gqf_rsqf_dump(gqf_rsqf cqf, void* host_ref_layout);
gqf_status gqf_rsqf_load(gqf_rsqf cqf, const void* host_ref_layout);

/* ============================================================ table images
 * A versioned byte image of a table, for files, wire transfers, fixtures and
 * snapshots (no reference counterpart: the reference only prints tables,
 * quotientFilter.cu:383-405, RSQF.cu:127-172).  A 64-byte header (magic
 * "GQFTABLE", version 1, layout, q, r, hash, payload size, FNV-1a 64 of the
 * payload, slots or blocks; little-endian) followed by one of the reference
 * layouts bit for bit:
 *   GQF_IMAGE_SQF_GPU  the GPU table bytes (quotientFilter.cu:151-170);
 *   GQF_IMAGE_SQF_CPU  the CPU_version's (r+3)-bit elements packed MSB-first
 *                      in u32 words (quotientFilter.cpp:114-130); it holds a
 *                      few slots fewer than the GPU table, saving a table that
 *                      uses them fails with GQF_ERR_CAPACITY;
 *   GQF_IMAGE_RSQF     packed reference blocks (RSQF.cuh:24-30).
 * An SQF image records the calling thread's SQF hash at save time; a load on
 * a thread whose hash differs fails with GQF_ERR_INVALID (the table would
 * answer that thread's queries with misses).  RSQF images record AP.  Loading
 * checks magic, version, geometry, hash and checksum before touching the
 * device. */
enum gqf_image_layout { GQF_IMAGE_SQF_GPU = 1, GQF_IMAGE_SQF_CPU = 2, GQF_IMAGE_RSQF = 3 };
/* Image size in bytes (0 for an unknown layout). */
size_t gqf_image_bytes(int layout, unsigned int q, unsigned int r);
gqf_status gqf_image_info(const void* image, size_t bytes, int* layout, unsigned int* q, unsigned int* r,
                          int* hash);
gqf_status gqf_sqf_image_save(gqf_sqf qf, int layout, void* image, size_t cap);
/* Either SQF layout into an initialised SQF of the same q and r. */
gqf_status gqf_sqf_image_load(gqf_sqf qf, const void* image, size_t bytes);
gqf_status gqf_rsqf_image_save(gqf_rsqf cqf, void* image, size_t cap);
gqf_status gqf_rsqf_image_load(gqf_rsqf cqf, const void* image, size_t bytes);

/* ================================================================ batches
 * Many SQFs per launch (no reference counterpart: the reference launches every
 * kernel per filter).  A batch binds F initialised SQFs of one r (any q), the
 * build keys of all of them back to back in d_keys (filter f owns
 * [key_offsets[f], key_offsets[f+1]); host array of F + 1 offsets) and their
 * lookup keys likewise (d_probe, probe_offsets), results to d_out in the
 * lookup layout.  Either key set may be absent (NULL offsets).  Creating the
 * batch autotunes its launch shapes on these inputs -- running the build
 * (which leaves every table built) and the lookups -- then captures both
 * pipelines as hipGraphs.  Buffers stay bound: refill them in place between
 * launches.  With a NULL stream the batch runs on a stream of its own that
 * waits for the null stream (graphs cannot be captured on the null stream).
 * Results equal the per-filter calls: gqf_sqf_bulk_build
 * (method irrelevant) with the same keys and no_dup, and unsorted
 * gqf_sqf_lookup; a filter whose keys do not fit keeps its table and reports
 * GQF_ERR_CAPACITY in filter_status. */
typedef struct gqf_sqf_batch_s* gqf_sqf_batch;

typedef struct gqf_batch_info {
  int filters;
  uint32_t build_keys, lookup_keys;
  int sort_passes, sort_chunks, sort_subtiles_per_chunk, rebuild_subtiles;
  int lookup_block, lookup_keys_per_thread;  /* chosen lookup shape */
  int build_graph_nodes, lookup_graph_nodes;
  float tune_build_ms[4];   /* build time per sort chunk size tried (1, 2, 4, 8 sub-tiles; 0 = skipped) */
  float tune_lookup_ms[15]; /* lookup time per shape tried: blocks 64..1024 x keys per thread 1, 2, 4 */
} gqf_batch_info;

gqf_status gqf_sqf_batch_create(gqf_sqf_batch* out, const gqf_sqf* filters, int count, int no_dup,
                                const uint32_t* d_keys, const uint64_t* key_offsets, const uint32_t* d_probe,
                                const uint64_t* probe_offsets, uint32_t* d_out, gqf_stream stream);
/* Rebuild every filter of the batch from its keys: one graph launch
 * (use_graph != 0) or the same kernels launched one by one.  filter_status
 * (F entries, may be NULL: then any failure returns GQF_ERR_CAPACITY). */
gqf_status gqf_sqf_batch_build(gqf_sqf_batch b, int use_graph, int* filter_status, float* ms);
/* All lookups of the batch (unsorted path): one graph launch, or the same
 * kernels one by one.  The launch first indexes every table as it is (the
 * occupied quotients and run starts before each 64-slot window), so a walk
 * through a long cluster finishes by binary searches; results equal the
 * per-filter walk, also after a filter was changed since the batch build. */
gqf_status gqf_sqf_batch_lookup(gqf_sqf_batch b, int use_graph, float* ms);
/* The same lookups in one launch, every walk run to its end through the
 * table (no window index): the per-filter calls' algorithm, batched.  The
 * results are the same; the time is the like-for-like comparison with the
 * per-filter calls. */
gqf_status gqf_sqf_batch_lookup_walk(gqf_sqf_batch b, float* ms);
gqf_status gqf_sqf_batch_info(gqf_sqf_batch b, gqf_batch_info* info);
gqf_status gqf_sqf_batch_destroy(gqf_sqf_batch b);

/* ================================================================ helpers
 * Device memory for callers without a HIP toolchain (ctypes and the like). */
gqf_status gqf_device_count(int* count);
gqf_status gqf_malloc(void** d_ptr, size_t bytes);
gqf_status gqf_free(void* d_ptr);
gqf_status gqf_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
gqf_status gqf_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
gqf_status gqf_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
gqf_status gqf_memset(void* d_ptr, int value, size_t bytes);
gqf_status gqf_synchronize(void);
/* A non-blocking stream of default priority for concurrent callers: calls on
 * different streams (from different host threads) run concurrently, each with
 * its own workspace; calls on one stream are serialised. */
gqf_status gqf_stream_create(gqf_stream* out);
/* Waits for the stream, releases its workspace and destroys it. */
gqf_status gqf_stream_destroy(gqf_stream s);
/* Workspace currently held for calls on the stream (grow-only until trimmed),
 * and its release. */
gqf_status gqf_stream_workspace_bytes(gqf_stream s, size_t* bytes);
gqf_status gqf_stream_trim(gqf_stream s);
/* What a workspace grows to when a call needs `bytes` more than it holds. */
size_t gqf_workspace_alloc_bytes(size_t bytes);

/* ================================================================ serving
 * A request loop over resident RSQFs (no reference counterpart; the
 * reference has one blocking call per batch).  Requests carry device key
 * buffers and a release time; the loop admits them against a device-memory
 * budget and executes them:
 *   coalesce != 0  one worker and stream per filter; a worker takes the
 *                  longest run of same-kind requests at the head of its
 *                  filter's queue (at most max_batch keys) and executes it as
 *                  one call: lookups gathered and their results scattered
 *                  back, inserts / deletes concatenated (the canonical layout
 *                  of M + A + B, and M - A - B = M - (A + B) as multisets).
 *                  Per filter, calls follow arrival order, so every result
 *                  and the end state equal executing the requests one by one;
 *                  different filters run concurrently on the device.
 *   coalesce == 0  plain execution: one call per request, in arrival order,
 *                  on one stream.
 * Admission: a call starts when the filter tables, the workspaces the
 * workers hold and the call's own need (workspace growth, gather buffers) fit
 * hbm_budget; otherwise it waits for running calls to finish, releasing the
 * workspaces of idle workers.  A request that cannot fit on its own is
 * rejected (status GQF_ERR_NOMEM) and the loop goes on. */
enum gqf_op { GQF_OP_LOOKUP = 0, GQF_OP_INSERT = 1, GQF_OP_DELETE = 2 };

typedef struct gqf_request {
  int filter;              /* index into the filter array */
  int op;                  /* gqf_op */
  int n;                   /* number of keys */
  const uint32_t* d_keys;  /* device keys */
  int32_t* d_out;          /* lookups: slot or -1 per key (device); else unused */
  double arrive_us;        /* release time in microseconds after the loop starts; non-decreasing */
  /* written by the loop */
  double start_us;         /* its call started (admitted) */
  double done_us;          /* its call completed */
  int status;              /* gqf_status of that call (GQF_ERR_NOMEM: rejected) */
  int call;                /* id of that call; coalesced requests share it */
} gqf_request;

typedef struct gqf_serve_config {
  size_t hbm_budget;       /* bytes: tables + held workspaces + gather buffers */
  int coalesce;            /* see above */
  int sorted_lookups;      /* lookups through the sorted entry point (the same walk) */
  int max_batch;           /* keys per coalesced call; 0 = 2^24 */
} gqf_serve_config;

typedef struct gqf_serve_stats {
  uint64_t calls;          /* device calls made */
  uint64_t requests;       /* requests served (not rejected) */
  uint64_t rejected;
  uint64_t keys;           /* keys of the served requests */
  double span_us;          /* first release to last completion */
  double p50_us, p99_us, mean_us;  /* latency: done_us - arrive_us over served requests */
  double mops;             /* keys / span */
  size_t peak_bytes;       /* highest accounted device bytes */
} gqf_serve_stats;

gqf_status gqf_rsqf_serve(gqf_rsqf* filters, int nfilters, gqf_request* reqs, int nreqs,
                          const gqf_serve_config* cfg, gqf_serve_stats* stats);

/* The same scheduler over caller-supplied operations (e.g. another filter
 * type, or a host model of the device for testing the scheduling itself). */
typedef struct gqf_serve_executor {
  void* user;
  /* device bytes worker `worker` must add to what it holds to run a call of
   * `op` on `filter` over n keys (coalesced: the call gathers its keys) */
  size_t (*grow_bytes)(void* user, int worker, int filter, int op, uint64_t n, int coalesced);
  /* bytes the worker holds now; release them (called only while it is idle) */
  size_t (*held_bytes)(void* user, int worker);
  void (*trim)(void* user, int worker);
  /* execute one call over the batch (same filter and op, arrival order);
   * returns a gqf_status */
  int (*run)(void* user, int worker, gqf_request** batch, int nbatch);
} gqf_serve_executor;

gqf_status gqf_serve_run(int nfilters, size_t resident_bytes, gqf_request* reqs, int nreqs,
                         const gqf_serve_config* cfg, const gqf_serve_executor* ex,
                         gqf_serve_stats* stats);

/* Random-access peak (no reference counterpart): n independent uniform-random
 * reads of aligned 64-byte sectors from a device buffer of buffer_bytes (one
 * sector per thread, plus a 4-byte index read and 4-byte result write per
 * thread, as a lookup's key and answer); best device time of reps timed runs
 * after one warm-up.  Lookups are measured against n / best_ms sectors/s. */
gqf_status gqf_gather_bench(size_t buffer_bytes, uint32_t n, int reps, float* best_ms);

/* Streaming peak (no reference counterpart): best device times of a read, a
 * copy and a write pass over buffers of `bytes` each (two are allocated),
 * after one warm-up round; 16-byte loads and stores, each workgroup over one
 * contiguous span.  Sized beyond the 256 MB Infinity Cache, the passes
 * measure HBM: read bytes / read_ms, 2 * bytes / copy_ms, bytes / write_ms. */
gqf_status gqf_stream_bench(size_t bytes, int reps, float* read_ms, float* copy_ms, float* write_ms);

/* ============================================================ kernel timing
 * Per-kernel device time measured with HIP events on the launch stream (two
 * events around every kernel launch while enabled; results accumulate per
 * kernel name when each call completes).  No reference counterpart: the
 * reference times whole calls only (e.g. quotientFilter.cu:279-310). */
gqf_status gqf_profile_enable(int on);
gqf_status gqf_profile_reset(void);
int gqf_profile_count(void);
/* total_bytes: algorithmic bytes (minimum HBM traffic of the launches' work,
 * as each launch site declares it); 0 for launches that declare none. */
gqf_status gqf_profile_get(int i, char* name, int name_cap, uint64_t* launches, double* total_ms,
                           double* total_bytes);

/* Host-side key stream of the reference drivers: init_genrand(seed) +
 * genrand_int32() (mt19937ar.cpp:57-70), i.e. std::mt19937, after skipping
 * `skip` outputs. */
void gqf_mt19937_fill(uint32_t seed, uint64_t skip, size_t n, uint32_t* host_out);

#ifdef __cplusplus
}
#endif
#endif /* GQF_H */
