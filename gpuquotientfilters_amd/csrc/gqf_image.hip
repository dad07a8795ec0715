// gqf_image.hip -- versioned table images: one byte format for files, wire
// transfers, fixtures and snapshots, carrying any of the three reference
// layouts bit for bit (host code only).
//
// Image = 64-byte header, then the payload:
//   0  char[8] magic "GQFTABLE"
//   8  u32 version (1)
//  12  u32 layout   1 SQF GPU table   calcNumSlotsGPU bytes of big-endian
//                                     (r+3)/8-byte elements (quotientFilter.cu:37-42, 151-170);
//                                     for (r+3)%8 != 0 the device words of
//                                     the packed table (gqf_sqf_table_bytes)
//                   2 SQF CPU table   (r+3)-bit elements packed MSB-first
//                                     into u32 words (quotientFilter.cpp:114-130)
//                   3 RSQF blocks     packed {u8 offset; u64 occupieds;
//                                     u64 runEnds; u64 remainders[r]} (RSQF.cuh:24-30)
//  16  u32 q, 20 u32 r
//  24  u32 hash     0 AP, 1 FNV: the hash the entries were made with
//  28  u32 reserved (0)
//  32  u64 payload bytes
//  40  u64 FNV-1a 64 of the payload
//  48  u64 slots (SQF) or blocks (RSQF) the payload holds
//  56  u64 reserved (0)
// All integers little-endian.  Loading checks magic, version, geometry,
// size and checksum before anything reaches the device.
#include "gqf_runtime.hpp"

#include <cstring>
#include <vector>

namespace gqf {
namespace {

constexpr char kMagic[8] = {'G', 'Q', 'F', 'T', 'A', 'B', 'L', 'E'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 64;

uint64_t fnv1a64(const uint8_t* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

void put32(uint8_t* p, uint32_t v) {
  for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
void put64(uint8_t* p, uint64_t v) {
  for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get32(const uint8_t* p) {
  uint32_t v = 0;
  for (int i = 0; i < 4; ++i) v |= (uint32_t)p[i] << (8 * i);
  return v;
}
uint64_t get64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
  return v;
}

uint32_t sqf_slots_of(unsigned q) { return gqf_sqf_num_slots(q); }

// Words of the CPU_version table (quotientFilter.cpp:60-70): (2^q (r+3) / 32
// words, one more when the remainder exceeds one, then * 1.1).
size_t cpu_words(unsigned q, unsigned r) {
  size_t slots = ((size_t)1 << q) * (r + 3) / 32;
  if (slots % 32 > 1) slots += 1;
  return (size_t)((double)slots * 1.1);
}

size_t payload_bytes(int layout, unsigned q, unsigned r) {
  if (layout == GQF_IMAGE_SQF_GPU) return gqf_sqf_table_bytes(q, r);
  if (layout == GQF_IMAGE_SQF_CPU) return cpu_words(q, r) * 4;
  if (layout == GQF_IMAGE_RSQF) return gqf_rsqf_ref_bytes(q, r);
  return 0;
}

bool is_packed(unsigned r) { return (r + 3) % 8 != 0; }

gqf_status check_sqf_geometry(const gqf_sqf& qf) {
  const unsigned r = qf.rbits;
  const bool ok_bpe = is_packed(r) ? (r + 3 <= 16 && qf.bytesPerElement == 0) : qf.bytesPerElement == (r + 3) / 8;
  if (!qf.table || qf.qbits < 1 || qf.qbits > 28 || r < 1 || qf.qbits + r > 32 || !ok_bpe) {
    set_error("not an initialised SQF");
    return GQF_ERR_INVALID;
  }
  return GQF_OK;
}

// Element of slot i of a GPU table held on the host: bpe big-endian bytes, or
// (bpe == 0) (r+3)-bit fields packed MSB-first in the table's u32 words.
uint32_t table_elem(const std::vector<uint8_t>& t, unsigned r, unsigned bpe, uint32_t i) {
  if (bpe) {
    uint32_t e = 0;
    for (unsigned k = 0; k < bpe; ++k) e = (e << 8) | t[(size_t)i * bpe + k];
    return e;
  }
  const unsigned w = r + 3;
  const uint64_t bit = (uint64_t)i * w;
  uint32_t e = 0;
  for (unsigned k = 0; k < w; ++k) {
    const uint64_t p = bit + k;
    uint32_t word;
    memcpy(&word, t.data() + 4 * (p / 32), 4);
    e = (e << 1) | ((word >> (31 - p % 32)) & 1u);
  }
  return e;
}
void table_set(std::vector<uint8_t>& t, unsigned r, unsigned bpe, uint32_t i, uint32_t e) {
  if (bpe) {
    for (unsigned k = 0; k < bpe; ++k) t[(size_t)i * bpe + k] = (uint8_t)(e >> (8 * (bpe - 1 - k)));
    return;
  }
  const unsigned w = r + 3;
  const uint64_t bit = (uint64_t)i * w;
  for (unsigned k = 0; k < w; ++k)
    if ((e >> (w - 1 - k)) & 1u) {
      const uint64_t p = bit + k;
      uint32_t word;
      memcpy(&word, t.data() + 4 * (p / 32), 4);
      word |= 1u << (31 - p % 32);
      memcpy(t.data() + 4 * (p / 32), &word, 4);
    }
}

void write_header(uint8_t* h, int layout, unsigned q, unsigned r, int hash, const uint8_t* payload, size_t pbytes,
                  uint64_t units) {
  memset(h, 0, kHeader);
  memcpy(h, kMagic, 8);
  put32(h + 8, kVersion);
  put32(h + 12, (uint32_t)layout);
  put32(h + 16, q);
  put32(h + 20, r);
  put32(h + 24, (uint32_t)hash);
  put64(h + 32, pbytes);
  put64(h + 40, fnv1a64(payload, pbytes));
  put64(h + 48, units);
}

// Validates an image; on success *layout, *q, *r, *hash and the payload.
gqf_status read_header(const void* image, size_t bytes, int* layout, unsigned* q, unsigned* r, int* hash,
                       const uint8_t** payload) {
  const uint8_t* h = static_cast<const uint8_t*>(image);
  if (!image || bytes < kHeader || memcmp(h, kMagic, 8) != 0) {
    set_error("not a table image (magic)");
    return GQF_ERR_INVALID;
  }
  if (get32(h + 8) != kVersion) {
    set_error("table image version %u is not supported (this library reads version %u)", get32(h + 8), kVersion);
    return GQF_ERR_UNSUPPORTED;
  }
  const int lay = (int)get32(h + 12);
  const unsigned qq = get32(h + 16), rr = get32(h + 20);
  const uint64_t pb = get64(h + 32);
  if (lay < GQF_IMAGE_SQF_GPU || lay > GQF_IMAGE_RSQF || qq > 30 || rr > 32 || get32(h + 24) > 1 ||
      pb != payload_bytes(lay, qq, rr) || bytes < kHeader + pb) {
    set_error("table image header is inconsistent (layout %d, q %u, r %u, %llu payload bytes)", lay, qq, rr,
              (unsigned long long)pb);
    return GQF_ERR_INVALID;
  }
  if (fnv1a64(h + kHeader, pb) != get64(h + 40)) {
    set_error("table image payload checksum mismatch");
    return GQF_ERR_INVALID;
  }
  *layout = lay;
  *q = qq;
  *r = rr;
  *hash = (int)get32(h + 24);
  *payload = h + kHeader;
  return GQF_OK;
}

}  // namespace
}  // namespace gqf

using namespace gqf;

extern "C" {

size_t gqf_image_bytes(int layout, unsigned int q, unsigned int r) {
  const size_t p = payload_bytes(layout, q, r);
  return p ? kHeader + p : 0;
}

gqf_status gqf_image_info(const void* image, size_t bytes, int* layout, unsigned int* q, unsigned int* r,
                          int* hash) {
  int lay = 0, hs = 0;
  unsigned qq = 0, rr = 0;
  const uint8_t* payload = nullptr;
  GQF_TRY(read_header(image, bytes, &lay, &qq, &rr, &hs, &payload));
  if (layout) *layout = lay;
  if (q) *q = qq;
  if (r) *r = rr;
  if (hash) *hash = hs;
  return GQF_OK;
}

gqf_status gqf_sqf_image_save(gqf_sqf qf, int layout, void* image, size_t cap) {
  GQF_TRY(check_sqf_geometry(qf));
  const unsigned q = qf.qbits, r = qf.rbits;
  if (layout != GQF_IMAGE_SQF_GPU && layout != GQF_IMAGE_SQF_CPU) {
    set_error("an SQF saves as layout GQF_IMAGE_SQF_GPU or GQF_IMAGE_SQF_CPU");
    return GQF_ERR_INVALID;
  }
  const size_t need = gqf_image_bytes(layout, q, r);
  if (!image || cap < need) {
    set_error("image buffer too small (%zu bytes needed)", need);
    return GQF_ERR_INVALID;
  }
  uint8_t* out = static_cast<uint8_t*>(image);
  uint8_t* payload = out + kHeader;
  const size_t tb = gqf_sqf_table_bytes(q, r);
  const uint32_t S = sqf_slots_of(q);
  if (layout == GQF_IMAGE_SQF_GPU) {
    GQF_HIP(hipMemcpy(payload, qf.table, tb, hipMemcpyDeviceToHost));
    write_header(out, layout, q, r, gqf_sqf_get_hash(), payload, tb, S);
    return GQF_OK;
  }
  // CPU layout: elements repacked MSB-first into (r+3)-bit fields of u32 words
  std::vector<uint8_t> t(tb);
  GQF_HIP(hipMemcpy(t.data(), qf.table, tb, hipMemcpyDeviceToHost));
  const size_t nw = cpu_words(q, r);
  const unsigned w = r + 3, bpe = qf.bytesPerElement;
  const uint64_t cap_slots = (uint64_t)nw * 32 / w;
  std::vector<uint32_t> words(nw, 0u);
  for (uint32_t i = 0; i < S; ++i) {
    const uint32_t e = table_elem(t, r, bpe, i);
    if (!e) continue;
    if (i >= cap_slots) {
      set_error("slot %u is used but the CPU layout holds %llu slots", i, (unsigned long long)cap_slots);
      return GQF_ERR_CAPACITY;
    }
    const uint64_t bit = (uint64_t)i * w;
    for (unsigned k = 0; k < w; ++k)
      if ((e >> (w - 1 - k)) & 1u) {
        const uint64_t p = bit + k;
        words[p / 32] |= 1u << (31 - p % 32);
      }
  }
  for (size_t i = 0; i < nw; ++i) put32(payload + 4 * i, words[i]);
  write_header(out, layout, q, r, gqf_sqf_get_hash(), payload, nw * 4, cap_slots);
  return GQF_OK;
}

gqf_status gqf_sqf_image_load(gqf_sqf qf, const void* image, size_t bytes) {
  GQF_TRY(check_sqf_geometry(qf));
  int lay = 0, hs = 0;
  unsigned q = 0, r = 0;
  const uint8_t* payload = nullptr;
  GQF_TRY(read_header(image, bytes, &lay, &q, &r, &hs, &payload));
  if (lay == GQF_IMAGE_RSQF || q != qf.qbits || r != qf.rbits) {
    set_error("image (layout %d, q %u, r %u) does not fit this SQF (q %u, r %u)", lay, q, r, qf.qbits, qf.rbits);
    return GQF_ERR_INVALID;
  }
  // the table's fingerprints were made with the image's hash; queries of the
  // calling thread use its own, so a mismatch would give silent misses
  if (hs != gqf_sqf_get_hash()) {
    set_error("image was saved with SQF hash %d but the calling thread uses hash %d (gqf_sqf_set_hash)", hs,
              gqf_sqf_get_hash());
    return GQF_ERR_INVALID;
  }
  const size_t tb = gqf_sqf_table_bytes(q, r);
  if (lay == GQF_IMAGE_SQF_GPU) {
    GQF_HIP(hipMemcpy(qf.table, payload, tb, hipMemcpyHostToDevice));
    return GQF_OK;
  }
  const size_t nw = cpu_words(q, r);
  const unsigned w = r + 3, bpe = qf.bytesPerElement;
  const uint32_t S = sqf_slots_of(q);
  const uint64_t cap_slots = std::min<uint64_t>((uint64_t)nw * 32 / w, S);
  std::vector<uint8_t> t(tb, 0);
  for (uint32_t i = 0; i < cap_slots; ++i) {
    const uint64_t bit = (uint64_t)i * w;
    uint32_t e = 0;
    for (unsigned k = 0; k < w; ++k) {
      const uint64_t p = bit + k;
      e = (e << 1) | ((get32(payload + 4 * (p / 32)) >> (31 - p % 32)) & 1u);
    }
    table_set(t, r, bpe, i, e);
  }
  GQF_HIP(hipMemcpy(qf.table, t.data(), tb, hipMemcpyHostToDevice));
  return GQF_OK;
}

gqf_status gqf_rsqf_image_save(gqf_rsqf cqf, void* image, size_t cap) {
  if (!cqf.blocks || cqf.numBlocks != gqf_rsqf_num_blocks(cqf.qbits)) {
    set_error("not an initialised RSQF");
    return GQF_ERR_INVALID;
  }
  const size_t need = gqf_image_bytes(GQF_IMAGE_RSQF, cqf.qbits, cqf.rbits);
  if (!image || cap < need) {
    set_error("image buffer too small (%zu bytes needed)", need);
    return GQF_ERR_INVALID;
  }
  uint8_t* out = static_cast<uint8_t*>(image);
  GQF_TRY(gqf_rsqf_dump(cqf, out + kHeader));
  write_header(out, GQF_IMAGE_RSQF, cqf.qbits, cqf.rbits, GQF_HASH_AP, out + kHeader, need - kHeader,
               cqf.numBlocks);
  return GQF_OK;
}

gqf_status gqf_rsqf_image_load(gqf_rsqf cqf, const void* image, size_t bytes) {
  int lay = 0, hs = 0;
  unsigned q = 0, r = 0;
  const uint8_t* payload = nullptr;
  GQF_TRY(read_header(image, bytes, &lay, &q, &r, &hs, &payload));
  if (lay != GQF_IMAGE_RSQF || q != cqf.qbits || r != cqf.rbits) {
    set_error("image (layout %d, q %u, r %u) does not fit this RSQF (q %u, r %u)", lay, q, r, cqf.qbits, cqf.rbits);
    return GQF_ERR_INVALID;
  }
  if (hs != GQF_HASH_AP) {
    set_error("RSQF image records hash %d; the RSQF hashes with AP only", hs);
    return GQF_ERR_INVALID;
  }
  return gqf_rsqf_load(cqf, payload);
}

}  // extern "C"
