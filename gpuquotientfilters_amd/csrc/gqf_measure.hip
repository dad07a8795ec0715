// gqf_measure.hip -- the two ceilings the kernels are measured against.
//
//   * Random-access peak: independent uniform-random reads of aligned 64-byte
//     sectors over a buffer the size of a filter table, one sector per
//     thread, as a lookup touches one (SQF) or two (RSQF) sectors per key.
//     Each thread also reads a 4-byte index and writes a 4-byte result, like
//     a lookup's key and answer.
//   * Streaming peak: sequential 16-byte-per-lane read, copy and write
//     kernels over buffers far larger than the 256 MB Infinity Cache, so
//     every pass goes to HBM.
#include "gqf_kernels.hpp"

namespace gqf {

// Sector of thread i: a bijective mix of i (fmix64) reduced to the sector count.
__global__ __launch_bounds__(256) void k_gather_sectors(const uint4* __restrict__ buf, uint64_t nsect,
                                                        const uint32_t* __restrict__ idx, uint32_t n,
                                                        uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = fmix64((uint64_t)idx[i] * 0x9E3779B97F4A7C15ull + 1) % nsect;
  const uint4* p = buf + 4 * s;
  const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
  out[i] = a.x ^ a.w ^ b.y ^ b.z ^ c.x ^ c.w ^ d.y ^ d.z;
}

// Streaming kernels: workgroup b owns the contiguous span [b*per, (b+1)*per)
// of 16-byte words; each lane issues UNR independent 16-byte loads before it
// uses any.  The read kernel folds what it reads into one word per thread so
// the loads are not dead.
constexpr int ST_THREADS = 256, ST_UNR = 8;
__global__ __launch_bounds__(ST_THREADS) void k_stream_read(const uint4* __restrict__ a, uint64_t n, uint64_t per,
                                                            uint32_t* __restrict__ sink) {
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  uint32_t acc = 0;
  uint64_t i = lo + threadIdx.x;
  // whole groups: no lane is predicated, so all UNR loads are in flight at once
  for (; i + (uint64_t)(ST_UNR - 1) * ST_THREADS < hi; i += (uint64_t)ST_THREADS * ST_UNR) {
    uint4 v[ST_UNR];
#pragma unroll
    for (int u = 0; u < ST_UNR; ++u) v[u] = a[i + (uint64_t)u * ST_THREADS];
#pragma unroll
    for (int u = 0; u < ST_UNR; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (; i < hi; i += ST_THREADS) {
    const uint4 v = a[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  sink[(uint64_t)blockIdx.x * ST_THREADS + threadIdx.x] = acc;
}
__global__ __launch_bounds__(ST_THREADS) void k_stream_copy(const uint4* __restrict__ a, uint4* __restrict__ b,
                                                            uint64_t n, uint64_t per) {
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  uint64_t i = lo + threadIdx.x;
  for (; i + (uint64_t)(ST_UNR - 1) * ST_THREADS < hi; i += (uint64_t)ST_THREADS * ST_UNR) {
    uint4 v[ST_UNR];
#pragma unroll
    for (int u = 0; u < ST_UNR; ++u) v[u] = a[i + (uint64_t)u * ST_THREADS];
#pragma unroll
    for (int u = 0; u < ST_UNR; ++u) b[i + (uint64_t)u * ST_THREADS] = v[u];
  }
  for (; i < hi; i += ST_THREADS) b[i] = a[i];
}
__global__ __launch_bounds__(ST_THREADS) void k_stream_write(uint4* __restrict__ b, uint64_t n, uint64_t per,
                                                             uint32_t v) {
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += ST_THREADS) b[i] = make_uint4(v, v ^ 1u, v ^ 2u, (uint32_t)i);
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ a, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = i;
}

}  // namespace gqf

using namespace gqf;

extern "C" {

gqf_status gqf_gather_bench(size_t buffer_bytes, uint32_t n, int reps, float* best_ms) {
  if (buffer_bytes < 64 || n == 0 || reps <= 0 || !best_ms) return GQF_ERR_INVALID;
  const uint64_t nsect = buffer_bytes / 64;
  uint4* buf = nullptr;
  uint32_t *idx = nullptr, *out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  gqf_status st = GQF_OK;
  auto fail = [&](hipError_t e, const char* what) {
    st = hip_fail(e, what, __FILE__, __LINE__);
  };
  hipError_t e;
  if ((e = hipMalloc((void**)&buf, nsect * 64)) != hipSuccess) fail(e, "hipMalloc");
  else if ((e = hipMalloc((void**)&idx, (size_t)n * 4)) != hipSuccess) fail(e, "hipMalloc");
  else if ((e = hipMalloc((void**)&out, (size_t)n * 4)) != hipSuccess) fail(e, "hipMalloc");
  else if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) fail(e, "hipEventCreate");
  if (st == GQF_OK) {
    (void)hipMemset(buf, 0x5A, nsect * 64);
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, nullptr, idx, n);
    float best = 1e30f;
    for (int r = 0; r <= reps && st == GQF_OK; ++r) {  // the first run warms the buffer into the caches
      (void)hipEventRecord(e0, nullptr);
      hipLaunchKernelGGL(k_gather_sectors, dim3((n + 255) / 256), dim3(256), 0, nullptr, buf, nsect, idx, n, out);
      (void)hipEventRecord(e1, nullptr);
      if ((e = hipEventSynchronize(e1)) != hipSuccess) {
        fail(e, "hipEventSynchronize");
        break;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) best = ms < best ? ms : best;
    }
    *best_ms = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  if (idx) (void)hipFree(idx);
  if (out) (void)hipFree(out);
  return st;
}

gqf_status gqf_stream_bench(size_t bytes, int reps, float* read_ms, float* copy_ms, float* write_ms) {
  if (bytes < (1u << 20) || reps <= 0 || !read_ms || !copy_ms || !write_ms) return GQF_ERR_INVALID;
  const uint64_t n = bytes / 16;
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  // 8 workgroups per CU, each over one contiguous span
  const uint32_t grid = (uint32_t)cus * 8;
  const uint64_t per = (n + grid - 1) / grid;
  uint4 *a = nullptr, *b = nullptr;
  uint32_t* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  gqf_status st = GQF_OK;
  hipError_t e;
  if ((e = hipMalloc((void**)&a, n * 16)) != hipSuccess) st = hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  else if ((e = hipMalloc((void**)&b, n * 16)) != hipSuccess) st = hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  else if ((e = hipMalloc((void**)&sink, (size_t)grid * ST_THREADS * 4)) != hipSuccess)
    st = hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  else if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
    st = hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
  float best[3] = {1e30f, 1e30f, 1e30f};
  if (st == GQF_OK) {
    hipLaunchKernelGGL(k_stream_write, dim3(grid), dim3(ST_THREADS), 0, nullptr, a, n, per, 0x5A5A5A5Au);
    hipLaunchKernelGGL(k_stream_write, dim3(grid), dim3(ST_THREADS), 0, nullptr, b, n, per, 0u);
    // kinds in turn, so no pass starts with the previous pass's tail in the caches
    for (int rep = 0; rep <= reps && st == GQF_OK; ++rep) {
      for (int kind = 0; kind < 3 && st == GQF_OK; ++kind) {
        (void)hipEventRecord(e0, nullptr);
        if (kind == 0) hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(ST_THREADS), 0, nullptr, a, n, per, sink);
        else if (kind == 1) hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(ST_THREADS), 0, nullptr, a, b, n, per);
        else hipLaunchKernelGGL(k_stream_write, dim3(grid), dim3(ST_THREADS), 0, nullptr, b, n, per, (uint32_t)rep);
        (void)hipEventRecord(e1, nullptr);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) {
          st = hip_fail(e, "hipEventSynchronize", __FILE__, __LINE__);
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best[kind]) best[kind] = ms;  // the first round warms up
      }
    }
    *read_ms = best[0];
    *copy_ms = best[1];
    *write_ms = best[2];
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (sink) (void)hipFree(sink);
  return st;
}

}  // extern "C"
