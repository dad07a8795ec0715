// gqf_runtime.hip -- host runtime of libgqf: errors, per-device state,
// workspace, the blocking/timing scaffold and the device-memory helpers of
// the C ABI.
#include "gqf_runtime.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace gqf {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

gqf_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return e == hipErrorOutOfMemory ? GQF_ERR_NOMEM : GQF_ERR_HIP;
}

std::mutex& api_mutex() {
  static std::mutex m;
  return m;
}

static DeviceState g_states[64];

gqf_status device_state(DeviceState** out) {
  int dev = 0;
  GQF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device index %d out of range", dev);
    return GQF_ERR_INVALID;
  }
  DeviceState& ds = g_states[dev];
  if (ds.device < 0) {
    GQF_HIP(hipMalloc((void**)&ds.counters, CTR_TOTAL * sizeof(uint32_t)));
    GQF_HIP(hipMemset(ds.counters, 0, CTR_TOTAL * sizeof(uint32_t)));
    GQF_HIP(hipHostMalloc((void**)&ds.h_pinned, CTR_TOTAL * sizeof(uint32_t), hipHostMallocDefault));
    GQF_HIP(hipEventCreate(&ds.ev0));
    GQF_HIP(hipEventCreate(&ds.ev1));
    ds.device = dev;
  }
  *out = &ds;
  return GQF_OK;
}

gqf_status ws_reserve(DeviceState& ds, size_t bytes) {
  if (bytes <= ds.ws_cap) return GQF_OK;
  size_t cap = bytes + bytes / 4 + (1u << 20);
  GQF_HIP(hipDeviceSynchronize());
  if (ds.ws) GQF_HIP(hipFree(ds.ws));
  ds.ws = nullptr;
  ds.ws_cap = 0;
  GQF_HIP(hipMalloc((void**)&ds.ws, cap));
  ds.ws_cap = cap;
  return GQF_OK;
}

gqf_status status_reserve(DeviceState& ds, size_t words) {
  if (words <= ds.status_cap) return GQF_OK;
  size_t cap = words + words / 4 + 4096;
  GQF_HIP(hipDeviceSynchronize());
  if (ds.status) GQF_HIP(hipFree(ds.status));
  ds.status = nullptr;
  ds.status_cap = 0;
  GQF_HIP(hipMalloc((void**)&ds.status, cap * sizeof(unsigned long long)));
  // Tagged look-back words must never match a live epoch by accident.
  GQF_HIP(hipMemset(ds.status, 0, cap * sizeof(unsigned long long)));
  GQF_HIP(hipDeviceSynchronize());
  ds.status_cap = cap;
  return GQF_OK;
}

gqf_status CallScope::begin(gqf_stream s) {
  GQF_TRY(device_state(&ds));
  stream = reinterpret_cast<hipStream_t>(s);
  GQF_HIP(hipEventRecord(ds->ev0, stream));
  return GQF_OK;
}

gqf_status CallScope::end(float* ms, bool check_device_err) {
  if (check_device_err) {
    GQF_HIP(hipMemcpyAsync(ds->h_pinned, ds->counters, CTR_TOTAL * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, stream));
  }
  GQF_HIP(hipEventRecord(ds->ev1, stream));
  GQF_HIP(hipEventSynchronize(ds->ev1));
  GQF_HIP(hipGetLastError());
  if (ms) {
    float t = 0.f;
    GQF_HIP(hipEventElapsedTime(&t, ds->ev0, ds->ev1));
    *ms = t;
  }
  if (check_device_err && (ds->h_pinned[CTR_ERR] & DEV_ERR_CAPACITY)) {
    set_error("capacity exceeded: the entries do not fit the table (table left unchanged)");
    return GQF_ERR_CAPACITY;
  }
  return GQF_OK;
}

}  // namespace gqf

using namespace gqf;

extern "C" {

const char* gqf_last_error(void) { return g_err; }

const char* gqf_status_string(gqf_status s) {
  switch (s) {
    case GQF_OK: return "ok";
    case GQF_ERR_INVALID: return "invalid argument";
    case GQF_ERR_CAPACITY: return "capacity exceeded";
    case GQF_ERR_HIP: return "HIP runtime error";
    case GQF_ERR_NOMEM: return "out of device memory";
    case GQF_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char* gqf_version(void) { return "gqf 0.1 gfx950"; }

gqf_status gqf_device_count(int* count) {
  if (!count) return GQF_ERR_INVALID;
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
    return GQF_ERR_HIP;
  }
  return GQF_OK;
}

gqf_status gqf_malloc(void** d_ptr, size_t bytes) {
  if (!d_ptr) return GQF_ERR_INVALID;
  GQF_HIP(hipMalloc(d_ptr, bytes ? bytes : 1));
  return GQF_OK;
}

gqf_status gqf_free(void* d_ptr) {
  if (d_ptr) GQF_HIP(hipFree(d_ptr));
  return GQF_OK;
}

gqf_status gqf_memcpy_htod(void* d_dst, const void* h_src, size_t bytes) {
  if (bytes) GQF_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return GQF_OK;
}

gqf_status gqf_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes) {
  if (bytes) GQF_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return GQF_OK;
}

gqf_status gqf_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes) {
  if (bytes) GQF_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
  return GQF_OK;
}

gqf_status gqf_memset(void* d_ptr, int value, size_t bytes) {
  if (bytes) GQF_HIP(hipMemset(d_ptr, value, bytes));
  GQF_HIP(hipDeviceSynchronize());
  return GQF_OK;
}

gqf_status gqf_synchronize(void) {
  GQF_HIP(hipDeviceSynchronize());
  return GQF_OK;
}

}  // extern "C"
