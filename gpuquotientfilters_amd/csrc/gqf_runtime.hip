// gqf_runtime.hip -- host runtime of libgqf: errors, per-device state,
// workspace, the blocking/timing scaffold and the device-memory helpers of
// the C ABI.
#include "gqf_runtime.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <atomic>
#include <cstdlib>
#include <vector>

namespace gqf {

static thread_local char g_err[512];
static thread_local int t_sqf_hash = GQF_HASH_AP;

uint32_t sqf_hash_flag() { return t_sqf_hash == GQF_HASH_FNV ? 0x100u : 0u; }
static thread_local int t_sqf_insert_path = GQF_INSERT_AUTO;
int sqf_insert_path() { return t_sqf_insert_path; }
static thread_local int t_rsqf_insert_path = GQF_INSERT_AUTO;
int rsqf_insert_path() { return t_rsqf_insert_path; }

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

// Contexts by (device, stream); the table only grows.
static std::mutex g_ctx_mu;
struct CtxEntry {
  int device;
  hipStream_t stream;
  DeviceState* ds;
};
static std::vector<CtxEntry>* g_ctx = nullptr;

gqf_status device_state(hipStream_t stream, DeviceState** out) {
  int dev = 0;
  GQF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (!g_ctx) g_ctx = new std::vector<CtxEntry>();
  for (const CtxEntry& e : *g_ctx)
    if (e.device == dev && e.stream == stream) {
      *out = e.ds;
      return GQF_OK;
    }
  DeviceState* ds = new DeviceState();
  GQF_HIP(hipMalloc((void**)&ds->counters, CTR_TOTAL * sizeof(uint32_t)));
  GQF_HIP(hipMemset(ds->counters, 0, CTR_TOTAL * sizeof(uint32_t)));
  GQF_HIP(hipHostMalloc((void**)&ds->h_pinned, CTR_TOTAL * sizeof(uint32_t), hipHostMallocDefault));
  GQF_HIP(hipEventCreate(&ds->ev0));
  GQF_HIP(hipEventCreate(&ds->ev1));
  GQF_HIP(hipHostMalloc((void**)&ds->h_crowd, 64, hipHostMallocMapped | hipHostMallocCoherent));
  GQF_HIP(hipHostGetDevicePointer((void**)&ds->d_crowd, ds->h_crowd, 0));
  GQF_HIP(hipEventCreateWithFlags(&ds->ev_crowd, hipEventDisableTiming));
  GQF_HIP(hipDeviceSynchronize());
  ds->device = dev;
  ds->stream = stream;
  g_ctx->push_back({dev, stream, ds});
  *out = ds;
  return GQF_OK;
}

static DeviceState* find_state(hipStream_t stream, int dev) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (g_ctx)
    for (const CtxEntry& e : *g_ctx)
      if (e.device == dev && e.stream == stream) return e.ds;
  return nullptr;
}

// Frees the context of `stream` on the current device (its stream is about
// to be destroyed); waits for the stream first.
static gqf_status release_state(hipStream_t stream) {
  int dev = 0;
  GQF_HIP(hipGetDevice(&dev));
  DeviceState* ds = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx) return GQF_OK;
    for (size_t i = 0; i < g_ctx->size(); ++i)
      if ((*g_ctx)[i].device == dev && (*g_ctx)[i].stream == stream) {
        ds = (*g_ctx)[i].ds;
        g_ctx->erase(g_ctx->begin() + (long)i);
        break;
      }
  }
  if (!ds) return GQF_OK;
  {
    std::lock_guard<std::mutex> lk(ds->mu);
    GQF_HIP(hipStreamSynchronize(stream));
    if (ds->ws) GQF_HIP(hipFree(ds->ws));
    GQF_HIP(hipFree(ds->counters));
    GQF_HIP(hipHostFree(ds->h_pinned));
    GQF_HIP(hipEventDestroy(ds->ev0));
    GQF_HIP(hipEventDestroy(ds->ev1));
    GQF_HIP(hipHostFree(ds->h_crowd));
    GQF_HIP(hipEventDestroy(ds->ev_crowd));
    for (size_t i = 0; i < ds->pool_size; ++i) (void)hipEventDestroy(ds->pool[i]);
    free(ds->pool);
    free(ds->pending);
  }
  delete ds;
  return GQF_OK;
}

size_t ws_alloc_bytes(size_t bytes) { return bytes + bytes / 4 + (1u << 20); }

gqf_status ws_reserve(DeviceState& ds, size_t bytes) {
  if (bytes <= ds.ws_cap) return GQF_OK;
  size_t cap = ws_alloc_bytes(bytes);
  GQF_HIP(hipStreamSynchronize(ds.stream));
  if (ds.ws) GQF_HIP(hipFree(ds.ws));
  ds.ws = nullptr;
  ds.ws_cap = 0;
  GQF_HIP(hipMalloc((void**)&ds.ws, cap));
  ds.ws_cap = cap;
  return GQF_OK;
}

// ---------------------------------------------------------------- kernel timing
struct ProfStat {
  std::string name;
  uint64_t launches;
  double ms;
  double bytes;
};
static std::atomic<bool> g_prof{false};
static std::mutex g_prof_mu;  // guards g_stats
static std::vector<ProfStat> g_stats;
static thread_local DeviceState* t_ctx = nullptr;  // context of the call running on this thread

bool prof_enabled() { return g_prof.load(std::memory_order_relaxed); }

static hipEvent_t pool_event(DeviceState& ds) {
  if (ds.pool_used == ds.pool_size) {
    const size_t n = ds.pool_size ? 2 * ds.pool_size : 256;
    hipEvent_t* p = static_cast<hipEvent_t*>(realloc(ds.pool, n * sizeof(hipEvent_t)));
    if (!p) return nullptr;
    ds.pool = p;
    for (size_t i = ds.pool_size; i < n; ++i)
      if (hipEventCreate(&ds.pool[i]) != hipSuccess) {
        ds.pool_size = i;
        return ds.pool_used < ds.pool_size ? ds.pool[ds.pool_used++] : nullptr;
      }
    ds.pool_size = n;
  }
  return ds.pool[ds.pool_used++];
}

ProfScope::ProfScope(const char* n, hipStream_t s, double algorithmic_bytes, const uint32_t* dev_count,
                     double bytes_per_count)
    : name(n), stream(s), on(prof_enabled() && t_ctx != nullptr), bytes(algorithmic_bytes), dcount(dev_count),
      dper(bytes_per_count) {
  if (on) {
    a = pool_event(*t_ctx);
    if (a) (void)hipEventRecord(a, stream);
    else on = false;
  }
}

ProfScope::~ProfScope() {
  if (!on) return;
  DeviceState& ds = *t_ctx;
  hipEvent_t b = pool_event(ds);
  if (!b) return;
  (void)hipEventRecord(b, stream);
  if (ds.pending_used == ds.pending_size) {
    const size_t n = ds.pending_size ? 2 * ds.pending_size : 256;
    ProfRec* p = static_cast<ProfRec*>(realloc(ds.pending, n * sizeof(ProfRec)));
    if (!p) return;
    ds.pending = p;
    ds.pending_size = n;
  }
  ds.pending[ds.pending_used++] = {name, a, b, bytes, dcount, dper};
}

void prof_resolve(DeviceState& ds) {
  if (ds.pending_used) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (size_t i = 0; i < ds.pending_used; ++i) {
      const ProfRec& r = ds.pending[i];
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
      double bytes = r.bytes;
      if (r.dcount) {  // the call has completed: its device counts are final
        uint32_t c = 0;
        if (hipMemcpy(&c, r.dcount, 4, hipMemcpyDeviceToHost) == hipSuccess) bytes += (double)c * r.dper;
      }
      std::string nm(r.name);
      while (!nm.empty() && (nm.front() == '(' || nm.front() == ' ')) nm.erase(nm.begin());
      while (!nm.empty() && (nm.back() == ')' || nm.back() == ' ')) nm.pop_back();
      bool found = false;
      for (ProfStat& st : g_stats)
        if (st.name == nm) {
          st.launches++;
          st.ms += ms;
          st.bytes += bytes;
          found = true;
          break;
        }
      if (!found) g_stats.push_back({nm, 1, (double)ms, bytes});
    }
  }
  ds.pending_used = 0;
  ds.pool_used = 0;
}

gqf_status CallScope::begin(gqf_stream s) {
  stream = reinterpret_cast<hipStream_t>(s);
  GQF_TRY(device_state(stream, &ds));
  lock = std::unique_lock<std::mutex>(ds->mu);
  t_ctx = ds;
  GQF_HIP(hipEventRecord(ds->ev0, stream));
  return GQF_OK;
}

CallScope::~CallScope() {
  if (ds && t_ctx == ds) {
    ds->pending_used = 0;  // a call that failed before end(): drop its timing records
    ds->pool_used = 0;
    t_ctx = nullptr;
  }
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
  prof_resolve(*ds);
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

gqf_status gqf_sqf_set_hash(int kind) {
  if (kind != GQF_HASH_AP && kind != GQF_HASH_FNV) {
    set_error("unknown hash kind %d (GQF_HASH_AP or GQF_HASH_FNV)", kind);
    return GQF_ERR_INVALID;
  }
  t_sqf_hash = kind;
  return GQF_OK;
}

int gqf_sqf_get_hash(void) { return t_sqf_hash; }

gqf_status gqf_sqf_set_insert_path(int path) {
  if (path != GQF_INSERT_AUTO && path != GQF_INSERT_REBUILD && path != GQF_INSERT_LOCAL) {
    set_error("unknown insert path %d (GQF_INSERT_AUTO, _REBUILD or _LOCAL)", path);
    return GQF_ERR_INVALID;
  }
  t_sqf_insert_path = path;
  return GQF_OK;
}

int gqf_sqf_get_insert_path(void) { return t_sqf_insert_path; }

gqf_status gqf_rsqf_set_insert_path(int path) {
  if (path != GQF_INSERT_AUTO && path != GQF_INSERT_REBUILD && path != GQF_INSERT_LOCAL) {
    set_error("unknown insert path %d (GQF_INSERT_AUTO, _REBUILD or _LOCAL)", path);
    return GQF_ERR_INVALID;
  }
  t_rsqf_insert_path = path;
  return GQF_OK;
}

int gqf_rsqf_get_insert_path(void) { return t_rsqf_insert_path; }

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

gqf_status gqf_profile_enable(int on) {
  g_prof.store(on != 0);
  return GQF_OK;
}

gqf_status gqf_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_stats.clear();
  return GQF_OK;
}

int gqf_profile_count(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return (int)g_stats.size();
}

gqf_status gqf_profile_get(int i, char* name, int name_cap, uint64_t* launches, double* total_ms,
                           double* total_bytes) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (i < 0 || i >= (int)g_stats.size()) return GQF_ERR_INVALID;
  const ProfStat& st = g_stats[(size_t)i];
  if (name && name_cap > 0) {
    snprintf(name, (size_t)name_cap, "%s", st.name.c_str());
  }
  if (launches) *launches = st.launches;
  if (total_ms) *total_ms = st.ms;
  if (total_bytes) *total_bytes = st.bytes;
  return GQF_OK;
}

void gqf_mt19937_fill(uint32_t seed, uint64_t skip, size_t n, uint32_t* host_out) {
  std::mt19937 g(seed);
  g.discard(skip);
  for (size_t i = 0; i < n; ++i) host_out[i] = (uint32_t)g();
}

gqf_status gqf_synchronize(void) {
  GQF_HIP(hipDeviceSynchronize());
  return GQF_OK;
}

gqf_status gqf_stream_create(gqf_stream* out) {
  if (!out) return GQF_ERR_INVALID;
  hipStream_t s = nullptr;
  GQF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = reinterpret_cast<gqf_stream>(s);
  return GQF_OK;
}

gqf_status gqf_stream_destroy(gqf_stream s) {
  if (s) {
    GQF_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(s)));
    GQF_TRY(release_state(reinterpret_cast<hipStream_t>(s)));
    GQF_HIP(hipStreamDestroy(reinterpret_cast<hipStream_t>(s)));
  }
  return GQF_OK;
}

gqf_status gqf_stream_workspace_bytes(gqf_stream s, size_t* bytes) {
  if (!bytes) return GQF_ERR_INVALID;
  int dev = 0;
  GQF_HIP(hipGetDevice(&dev));
  DeviceState* ds = find_state(reinterpret_cast<hipStream_t>(s), dev);
  *bytes = 0;
  if (ds) {
    std::lock_guard<std::mutex> lk(ds->mu);
    *bytes = ds->ws_cap;
  }
  return GQF_OK;
}

gqf_status gqf_stream_trim(gqf_stream s) {
  int dev = 0;
  GQF_HIP(hipGetDevice(&dev));
  DeviceState* ds = find_state(reinterpret_cast<hipStream_t>(s), dev);
  if (!ds) return GQF_OK;
  std::lock_guard<std::mutex> lk(ds->mu);
  GQF_HIP(hipStreamSynchronize(ds->stream));
  if (ds->ws) GQF_HIP(hipFree(ds->ws));
  ds->ws = nullptr;
  ds->ws_cap = 0;
  return GQF_OK;
}

size_t gqf_workspace_alloc_bytes(size_t bytes) { return bytes ? ws_alloc_bytes(bytes) : 0; }

}  // extern "C"
