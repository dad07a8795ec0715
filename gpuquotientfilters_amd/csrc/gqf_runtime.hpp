// gqf_runtime.hpp -- host-side runtime of libgqf: error reporting, the
// per-device workspace arena, and the stream/timing scaffold shared by every
// entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "../../include/gqf.h"

namespace gqf {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
gqf_status hip_fail(hipError_t e, const char* what, const char* file, int line);

#define GQF_HIP(x)                                                      \
  do {                                                                  \
    hipError_t _e = (x);                                                \
    if (_e != hipSuccess) return ::gqf::hip_fail(_e, #x, __FILE__, __LINE__); \
  } while (0)

#define GQF_TRY(x)                      \
  do {                                  \
    gqf_status _s = (x);                \
    if (_s != GQF_OK) return _s;        \
  } while (0)

// SQF hash of the calling thread (gqf_sqf_set_hash) as a fingerprint-width
// flag: 0 (AP) or FP_FNV.
uint32_t sqf_hash_flag();
// SQF insert path of the calling thread (gqf_sqf_set_insert_path).
int sqf_insert_path();
int rsqf_insert_path();

// ---------------------------------------------------------------- device state
// Small device counters used by the kernels (index constants below).
enum Counter : int {
  CTR_OP = 8,          // first counter reset by every rebuild
  CTR_COUNT = 9,       // number of entries kept by the location scan
  CTR_ERR = 10,        // error flags raised on the device (bit 0: capacity)
  CTR_MISC = 11,       // 5 general-purpose slots
  CTR_TOTAL = 16
};
constexpr uint32_t DEV_ERR_CAPACITY = 1u;
// Per-call entry limit (scan states hold 32-bit counts and locations; the
// merged table + batch of an insert must stay below it).
constexpr uint64_t GQF_MAX_ITEMS = (1ull << 30) - 4096;

struct ProfRec {
  const char* name;
  hipEvent_t a, b;
  double bytes;
  const uint32_t* dcount;  // device count whose value adds dper bytes each (may be NULL)
  double dper;
};

// Execution context of one (device, stream) pair: workspace arena, device
// counters and their pinned mirror, timing events.  Calls on the same stream
// are serialized by the context's mutex; calls on different streams run
// concurrently (each with its own workspace).
struct DeviceState {
  int device = -1;
  hipStream_t stream = nullptr;
  std::mutex mu;
  char* ws = nullptr;               // workspace arena (grow-only)
  size_t ws_cap = 0;
  uint32_t* counters = nullptr;     // CTR_TOTAL device counters
  uint32_t* h_pinned = nullptr;     // pinned host mirror of the counters
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // sort probe: a word of host-coherent pinned memory (host and device views)
  // that a sort's row scan raises when its top-digit segments are too large
  // for the segment sort, and the event recorded after that scan
  uint32_t* h_crowd = nullptr;
  uint32_t* d_crowd = nullptr;
  hipEvent_t ev_crowd = nullptr;
  // after a crowded sort the next crowd_skip sorts take the chunked passes
  // directly; crowd_backoff doubles with every crowded probe in a row.  One
  // pair per key width: 32-bit (SQF, AP hash) and 64-bit (wide) fingerprints
  // crowd differently.
  uint32_t crowd_skip[2] = {0, 0}, crowd_backoff[2] = {0, 0};
  // per-launch timing (gqf_profile_enable): event pool and launches to resolve
  hipEvent_t* pool = nullptr;
  size_t pool_size = 0, pool_used = 0;
  ProfRec* pending = nullptr;
  size_t pending_size = 0, pending_used = 0;
};

// The context of the calling thread's current device and `stream`, created
// on first use (contexts live until the process ends).
gqf_status device_state(hipStream_t stream, DeviceState** out);
// Bytes a workspace grows to when a call needs `bytes` more than it holds.
size_t ws_alloc_bytes(size_t bytes);
gqf_status ws_reserve(DeviceState& ds, size_t bytes);

// Bump allocator over the workspace of one call.
struct Arena {
  char* base;
  size_t cap;
  size_t used = 0;
  Arena(char* b, size_t c) : base(b), cap(c) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    used = off + count * sizeof(T);
    return reinterpret_cast<T*>(base + off);
  }
};
__host__ __device__ inline size_t aligned_bytes(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// ---------------------------------------------------------------- kernel timing
// When enabled (gqf_profile_enable), every launch is bracketed by two events
// on its stream; elapsed times are accumulated per kernel name when the
// enclosing call completes.
bool prof_enabled();
struct ProfScope {
  const char* name;
  hipStream_t stream;
  hipEvent_t a = nullptr;
  bool on;
  double bytes;
  const uint32_t* dcount;
  double dper;
  ProfScope(const char* n, hipStream_t s, double algorithmic_bytes = 0.0, const uint32_t* dev_count = nullptr,
            double bytes_per_count = 0.0);
  ~ProfScope();
};
void prof_resolve(DeviceState& ds);

#define GQF_LAUNCH(kernel, grid, block, shm, stream, ...)              \
  do {                                                                 \
    ::gqf::ProfScope _gqf_ps(#kernel, (stream));                       \
    hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__); \
  } while (0)

// Same, recording the launch's algorithmic bytes (minimum HBM traffic of the
// work it does) for roofline accounting.
#define GQF_LAUNCH_B(bytes, kernel, grid, block, shm, stream, ...)     \
  do {                                                                 \
    ::gqf::ProfScope _gqf_ps(#kernel, (stream), (double)(bytes));      \
    hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__); \
  } while (0)

// Same, where part of the bytes depends on a count that only the device
// knows at launch time: bytes + (*dev_count) * per_count, read back when the
// call's timings are resolved.
#define GQF_LAUNCH_BC(bytes, dev_count, per_count, kernel, grid, block, shm, stream, ...)       \
  do {                                                                                        \
    ::gqf::ProfScope _gqf_ps(#kernel, (stream), (double)(bytes), (dev_count), (double)(per_count)); \
    hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);                        \
  } while (0)

// Timing/blocking scaffold of every entry point: takes the stream's context
// (and its lock), records the start event; end() records the stop event,
// waits, reads the device error counter and reports ms.
struct CallScope {
  DeviceState* ds = nullptr;
  hipStream_t stream = nullptr;
  std::unique_lock<std::mutex> lock;
  gqf_status begin(gqf_stream s);
  // Waits for the call's work; reports the device time since begin() and, when
  // check_device_err, a capacity error raised by the kernels.
  gqf_status end(float* ms, bool check_device_err);
  ~CallScope();
};

}  // namespace gqf
