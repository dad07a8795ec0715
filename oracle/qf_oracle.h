/* qf_oracle.h -- CPU oracle for the quotient-filter engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Contents (each function cites the reference code it restates):
 *   - hashes: Normal_APHash (+ the project's wide fingerprint for q+r > 32),
 *     FNV-1 (CPU_version hash);
 *   - SQF byte-slot codec (GPU_version layout) and packed-word codec
 *     (CPU_version layout);
 *   - SQF canonical layout builder (closed form) and faithful sequential
 *     restatements of insert / lookup / delete / fingerprint extraction;
 *   - RSQF packed-block codec (reference layout, runtime r), rank/select,
 *     faithful lookup and insert (including the findFirstUnusedSlot quirk),
 *     and the canonical RSQF builder;
 *   - std::mt19937 key streams and CPU-baseline timing loops.
 */
#ifndef QF_ORACLE_H
#define QF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NOT_FOUND 0xFFFFFFFFu

/* status codes shared with the product library's meaning */
#define ORC_OK 0
#define ORC_ERR_INVALID 1
#define ORC_ERR_CAPACITY 2

/* ---------------- keys ---------------- */
void orc_mt19937(uint32_t seed, uint64_t skip, size_t n, uint32_t* out);

/* ---------------- hashes ---------------- */
uint32_t orc_ap_rounds(uint32_t key, uint32_t state);
uint64_t orc_fmix64(uint64_t k);
uint64_t orc_fp_ap(uint32_t key, unsigned q, unsigned r);
uint64_t orc_fp_fnv(uint32_t key, unsigned q, unsigned r);
/* hash_kind: 0 = AP (wide rule above 32 bits), 1 = FNV */
void orc_fingerprints(const uint32_t* keys, size_t n, unsigned q, unsigned r, int hash_kind,
                      uint64_t* out);

/* ---------------- SQF geometry ---------------- */
uint32_t orc_sqf_num_slots(unsigned q);
size_t orc_sqf_table_bytes(unsigned q, unsigned r);
size_t orc_sqf_cpu_words(unsigned q, unsigned r);

/* ---------------- SQF codecs ---------------- */
uint32_t orc_sqf_get(const uint8_t* table, unsigned r, uint32_t slot);
void orc_sqf_set(uint8_t* table, unsigned r, uint32_t slot, uint32_t elem);
void orc_sqf_decode_bytes(const uint8_t* table, unsigned r, uint32_t nslots, uint32_t* elems);
void orc_sqf_decode_words(const uint32_t* words, size_t nwords, unsigned r, uint32_t nslots,
                          uint32_t* elems);

/* ---------------- SQF canonical layout ---------------- */
/* fps: any order; sorts a private copy.  Returns ORC_OK or ORC_ERR_CAPACITY.
 * table must hold orc_sqf_table_bytes(q, r) bytes; it is fully overwritten. */
int orc_sqf_canon(const uint64_t* fps, size_t n, unsigned q, unsigned r, int no_dup,
                  uint8_t* table);

/* ---------------- SQF faithful sequential restatements ---------------- */
uint32_t orc_sqf_lookup_fp(const uint8_t* table, unsigned q, unsigned r, uint64_t fp);
/* returns slot, or ORC_NOT_FOUND; *status = ORC_ERR_CAPACITY on tail overflow */
uint32_t orc_sqf_insert_fp(uint8_t* table, unsigned q, unsigned r, uint64_t fp, int* status);
/* returns 1 if a copy was removed, 0 otherwise */
int orc_sqf_delete_fp(uint8_t* table, unsigned q, unsigned r, uint64_t fp);
/* extract every stored fingerprint in slot order; returns count */
size_t orc_sqf_extract(const uint8_t* table, unsigned q, unsigned r, uint64_t* out);

void orc_sqf_lookup_keys(const uint8_t* table, unsigned q, unsigned r, int hash_kind,
                         const uint32_t* keys, size_t n, uint32_t* out);
int orc_sqf_insert_keys(uint8_t* table, unsigned q, unsigned r, int hash_kind,
                        const uint32_t* keys, size_t n);
size_t orc_sqf_delete_keys(uint8_t* table, unsigned q, unsigned r, int hash_kind,
                           const uint32_t* keys, size_t n);

/* ---------------- RSQF ---------------- */
uint32_t orc_rsqf_num_blocks(unsigned q);
size_t orc_rsqf_block_bytes(unsigned r);
size_t orc_rsqf_table_bytes(unsigned q, unsigned r);
uint32_t orc_rank(uint64_t x, unsigned i);
uint32_t orc_select(uint64_t x, unsigned k);
uint32_t orc_rsqf_get_rem(const uint8_t* table, unsigned r, uint32_t block, unsigned slot);

int orc_rsqf_canon(const uint64_t* fps, size_t n, unsigned q, unsigned r, uint8_t* table);
uint32_t orc_rsqf_lookup_fp(const uint8_t* table, unsigned q, unsigned r, uint64_t fp);
uint32_t orc_rsqf_insert_fp(uint8_t* table, unsigned q, unsigned r, uint64_t fp);
void orc_rsqf_lookup_keys(const uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                          size_t n, uint32_t* out);
size_t orc_rsqf_insert_keys(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                            size_t n, uint32_t* out);
/* canonical decode: every stored fingerprint (ascending) of a canonical table */
size_t orc_rsqf_extract(const uint8_t* table, unsigned q, unsigned r, uint64_t* out);

/* ---------------- CPU baseline loops (steady_clock seconds) ---------------- */
double orc_time_sqf_insert(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys, size_t n);
double orc_time_sqf_lookup(const uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                           size_t n, uint32_t* out, int threads);
double orc_time_sqf_delete(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys, size_t n);
double orc_time_rsqf_insert(uint8_t* table, unsigned q, unsigned r, const uint32_t* keys, size_t n);
double orc_time_rsqf_lookup(const uint8_t* table, unsigned q, unsigned r, const uint32_t* keys,
                            size_t n, uint32_t* out, int threads);

#ifdef __cplusplus
}
#endif
#endif
