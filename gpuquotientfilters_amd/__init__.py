"""MI355X-native quotient filters (SQF + RSQF) -- Python binding of libgqf.so.

The library is a C ABI (``include/gqf.h``) over hand-written HIP kernels for
gfx950.  This module binds it with ctypes and mirrors the reference host API
names so scripts read like the reference drivers:

    qf = SQF.initFilterGPU(q, r)                 # quotientFilter.cuh:33
    ms = qf.bulkBuildSegmentedLayouts(d_keys, n, NoDuplicates=True)
    ms = qf.launchUnsortedLookups(d_keys, n, d_out)
    ms = qf.superclusterDeletes(d_keys, n)

    cqf = RSQF.initCQFGPU(q, r)                  # RSQF.cuh:41 (r is runtime here)
    ms = cqf.insertGPU(d_keys, n, d_out)
    ms = cqf.launchLookups(d_keys, n, d_out)     # sorted

Device memory is a :class:`DeviceArray` (library-allocated, numpy round
trips).  Nothing here computes on the CPU: if ``libgqf.so`` is missing the
import of any entry point raises :class:`GQFError`.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgqf.so")

NOT_FOUND = 0xFFFFFFFF
NOALLOC_MAX_KEYS = 1 << 21  # GQF_NOALLOC_MAX_KEYS
OK, ERR_INVALID, ERR_CAPACITY, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED = range(6)
BUILD_SEGMENTED_LAYOUTS, BUILD_PARALLEL_MERGING, BUILD_SEQUENTIAL_SHIFTS = 0, 1, 2


class GQFError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"gqf status {status}: {msg}")
        self.status = status


class CapacityError(GQFError):
    pass


class _SQF(C.Structure):
    _fields_ = [("qbits", C.c_uint), ("rbits", C.c_uint), ("bytesPerElement", C.c_uint), ("table", C.c_void_p)]


class _RSQF(C.Structure):
    _fields_ = [("qbits", C.c_uint), ("numBlocks", C.c_uint), ("blocks", C.c_void_p), ("rbits", C.c_uint)]


OP_LOOKUP, OP_INSERT, OP_DELETE = 0, 1, 2
HASH_AP, HASH_FNV = 0, 1
INSERT_AUTO, INSERT_REBUILD, INSERT_LOCAL = 0, 1, 2
IMAGE_SQF_GPU, IMAGE_SQF_CPU, IMAGE_RSQF = 1, 2, 3


class Request(C.Structure):
    """gqf_request (gqf.h): one served request."""
    _fields_ = [("filter", C.c_int), ("op", C.c_int), ("n", C.c_int), ("d_keys", C.c_void_p),
                ("d_out", C.c_void_p), ("arrive_us", C.c_double), ("start_us", C.c_double),
                ("done_us", C.c_double), ("status", C.c_int), ("call", C.c_int)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("coalesce", C.c_int), ("sorted_lookups", C.c_int),
                ("max_batch", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("calls", C.c_uint64), ("requests", C.c_uint64), ("rejected", C.c_uint64), ("keys", C.c_uint64),
                ("span_us", C.c_double), ("p50_us", C.c_double), ("p99_us", C.c_double), ("mean_us", C.c_double),
                ("mops", C.c_double), ("peak_bytes", C.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


GROW_FN = C.CFUNCTYPE(C.c_size_t, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int)
HELD_FN = C.CFUNCTYPE(C.c_size_t, C.c_void_p, C.c_int)
TRIM_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int)
RUN_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.POINTER(Request)), C.c_int)


class ServeExecutor(C.Structure):
    _fields_ = [("user", C.c_void_p), ("grow_bytes", GROW_FN), ("held_bytes", HELD_FN), ("trim", TRIM_FN),
                ("run", RUN_FN)]


class BatchInfo(C.Structure):
    """gqf_batch_info (gqf.h): the plan a batch chose."""
    _fields_ = [("filters", C.c_int), ("build_keys", C.c_uint32), ("lookup_keys", C.c_uint32),
                ("sort_passes", C.c_int), ("sort_chunks", C.c_int), ("sort_subtiles_per_chunk", C.c_int),
                ("rebuild_subtiles", C.c_int), ("lookup_block", C.c_int), ("lookup_keys_per_thread", C.c_int),
                ("build_graph_nodes", C.c_int), ("lookup_graph_nodes", C.c_int),
                ("tune_build_ms", C.c_float * 4), ("tune_lookup_ms", C.c_float * 15)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_[:11]}
        d["tune_build_ms"] = [round(x, 4) for x in self.tune_build_ms if x > 0]
        d["tune_lookup_ms"] = [round(x, 4) for x in self.tune_lookup_ms if x > 0]
        return d


_st = C.c_int
_vp = C.c_void_p
_f = C.POINTER(C.c_float)
_SIGS = {
    "gqf_last_error": (C.c_char_p, []),
    "gqf_status_string": (C.c_char_p, [_st]),
    "gqf_version": (C.c_char_p, []),
    "gqf_device_count": (_st, [C.POINTER(C.c_int)]),
    "gqf_malloc": (_st, [C.POINTER(_vp), C.c_size_t]),
    "gqf_free": (_st, [_vp]),
    "gqf_memcpy_htod": (_st, [_vp, _vp, C.c_size_t]),
    "gqf_memcpy_dtoh": (_st, [_vp, _vp, C.c_size_t]),
    "gqf_memcpy_dtod": (_st, [_vp, _vp, C.c_size_t]),
    "gqf_memset": (_st, [_vp, C.c_int, C.c_size_t]),
    "gqf_synchronize": (_st, []),
    "gqf_image_bytes": (C.c_size_t, [C.c_int, C.c_uint, C.c_uint]),
    "gqf_image_info": (_st, [_vp, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                             C.POINTER(C.c_int)]),
    "gqf_sqf_image_save": (_st, [_SQF, C.c_int, _vp, C.c_size_t]),
    "gqf_sqf_image_load": (_st, [_SQF, _vp, C.c_size_t]),
    "gqf_rsqf_image_save": (_st, [_RSQF, _vp, C.c_size_t]),
    "gqf_rsqf_image_load": (_st, [_RSQF, _vp, C.c_size_t]),
    "gqf_sqf_set_hash": (_st, [C.c_int]),
    "gqf_sqf_get_hash": (C.c_int, []),
    "gqf_sqf_set_insert_path": (_st, [C.c_int]),
    "gqf_sqf_get_insert_path": (C.c_int, []),
    "gqf_rsqf_set_insert_path": (_st, [C.c_int]),
    "gqf_rsqf_get_insert_path": (C.c_int, []),
    "gqf_stream_create": (_st, [C.POINTER(_vp)]),
    "gqf_stream_destroy": (_st, [_vp]),
    "gqf_stream_workspace_bytes": (_st, [_vp, C.POINTER(C.c_size_t)]),
    "gqf_stream_trim": (_st, [_vp]),
    "gqf_workspace_alloc_bytes": (C.c_size_t, [C.c_size_t]),
    "gqf_rsqf_workspace_bytes": (C.c_size_t, [_RSQF, C.c_int, C.c_int, C.c_int]),
    "gqf_rsqf_serve": (_st, [C.POINTER(_RSQF), C.c_int, C.POINTER(Request), C.c_int, C.POINTER(ServeConfig),
                             C.POINTER(ServeStats)]),
    "gqf_sqf_batch_create": (_st, [C.POINTER(_vp), C.POINTER(_SQF), C.c_int, C.c_int, _vp, C.POINTER(C.c_uint64),
                                   _vp, C.POINTER(C.c_uint64), _vp, _vp]),
    "gqf_sqf_batch_build": (_st, [_vp, C.c_int, C.POINTER(C.c_int), _f]),
    "gqf_sqf_batch_lookup": (_st, [_vp, C.c_int, _f]),
    "gqf_sqf_batch_lookup_walk": (_st, [_vp, _f]),
    "gqf_sqf_batch_info": (_st, [_vp, C.POINTER(BatchInfo)]),
    "gqf_sqf_batch_destroy": (_st, [_vp]),
    "gqf_serve_run": (_st, [C.c_int, C.c_size_t, C.POINTER(Request), C.c_int, C.POINTER(ServeConfig),
                            C.POINTER(ServeExecutor), C.POINTER(ServeStats)]),
    "gqf_profile_enable": (_st, [C.c_int]),
    "gqf_profile_reset": (_st, []),
    "gqf_profile_count": (C.c_int, []),
    "gqf_profile_get": (_st, [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]),
    "gqf_mt19937_fill": (None, [C.c_uint32, C.c_uint64, C.c_size_t, _vp]),
    "gqf_gather_bench": (_st, [C.c_size_t, C.c_uint32, C.c_int, _f]),
    "gqf_stream_bench": (_st, [C.c_size_t, C.c_int, _f, _f, _f]),
    "gqf_sqf_table_bytes": (C.c_size_t, [C.c_uint, C.c_uint]),
    "gqf_sqf_num_slots": (C.c_uint32, [C.c_uint]),
    "gqf_sqf_init": (_st, [C.POINTER(_SQF), C.c_uint, C.c_uint]),
    "gqf_sqf_destroy": (_st, [C.POINTER(_SQF)]),
    "gqf_sqf_bulk_build": (_st, [_SQF, C.c_int, C.c_int, _vp, C.c_int, _vp, _f]),
    "gqf_sqf_bulk_build_noalloc": (_st, [_SQF, C.c_int, C.c_int, _vp, C.c_int, _vp, _f]),
    "gqf_sqf_insert": (_st, [_SQF, C.c_int, _vp, _vp, _f]),
    "gqf_sqf_insert_via_merge": (_st, [_SQF, _vp, C.c_int, _vp, C.c_int, C.c_int, _vp, _f]),
    "gqf_sqf_lookup": (_st, [_SQF, C.c_int, C.c_int, _vp, _vp, _vp, _f]),
    "gqf_sqf_delete": (_st, [_SQF, C.c_int, _vp, _vp, _f]),
    "gqf_sqf_merge": (_st, [_SQF, _SQF, C.c_int, _vp, _f]),
    "gqf_rsqf_num_blocks": (C.c_uint32, [C.c_uint]),
    "gqf_rsqf_table_bytes": (C.c_size_t, [C.c_uint, C.c_uint]),
    "gqf_rsqf_ref_bytes": (C.c_size_t, [C.c_uint, C.c_uint]),
    "gqf_rsqf_init": (_st, [C.POINTER(_RSQF), C.c_uint, C.c_uint]),
    "gqf_rsqf_destroy": (_st, [C.POINTER(_RSQF)]),
    "gqf_rsqf_bulk_build": (_st, [_RSQF, C.c_int, _vp, _vp, _f]),
    "gqf_rsqf_insert": (_st, [_RSQF, C.c_int, _vp, _vp, _vp, _f]),
    "gqf_rsqf_insert_report": (_st, [_RSQF, C.c_int, _vp, _vp, C.POINTER(C.c_uint32), _vp, _f]),
    "gqf_rsqf_lookup": (_st, [_RSQF, C.c_int, C.c_int, _vp, _vp, _vp, _f]),
    "gqf_rsqf_delete": (_st, [_RSQF, C.c_int, _vp, _vp, _f]),
    "gqf_rsqf_count": (_st, [_RSQF, C.c_int, _vp, _vp, _vp, _f]),
    "gqf_rsqf_dump": (_st, [_RSQF, _vp]),
    "gqf_rsqf_load": (_st, [_RSQF, _vp]),
}
EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """The loaded libgqf.so (raises GQFError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GQFError(ERR_UNSUPPORTED, f"{LIB_PATH} not built: run `python -m gpuquotientfilters_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != OK:
        msg = lib().gqf_last_error().decode(errors="replace")
        if status == ERR_CAPACITY:
            raise CapacityError(status, msg)
        raise GQFError(status, msg)


def device_count():
    n = C.c_int(0)
    st = lib().gqf_device_count(C.byref(n))
    return n.value if st == OK else 0


def synchronize():
    check(lib().gqf_synchronize())


def profile_enable(on=True):
    check(lib().gqf_profile_enable(int(bool(on))))


def profile_reset():
    check(lib().gqf_profile_reset())


def profile_stats():
    """{kernel name: (launches, total ms, total algorithmic bytes)} accumulated
    while profiling was enabled."""
    out = {}
    buf = C.create_string_buffer(512)
    for i in range(lib().gqf_profile_count()):
        n, ms, by = C.c_uint64(0), C.c_double(0), C.c_double(0)
        check(lib().gqf_profile_get(i, buf, 512, C.byref(n), C.byref(ms), C.byref(by)))
        out[buf.value.decode()] = (n.value, ms.value, by.value)
    return out


def set_sqf_hash(kind):
    """Hash of this thread's SQF calls: HASH_AP (default, the reference GPU
    code's) or HASH_FNV (the reference CPU_version's)."""
    check(lib().gqf_sqf_set_hash(int(kind)))


def set_sqf_insert_path(path):
    """How this thread's SQF inserts update the table: INSERT_AUTO (default),
    INSERT_REBUILD (every slot) or INSERT_LOCAL (only the regions a batch
    changes).  The resulting tables are the same."""
    check(lib().gqf_sqf_set_insert_path(int(path)))


def sqf_insert_path():
    return lib().gqf_sqf_get_insert_path()


def set_rsqf_insert_path(path):
    """How this thread's RSQF inserts update the table: INSERT_AUTO (default)
    or INSERT_LOCAL (in place, window by window, falling back to the rebuild
    where a window cannot be rewritten alone), INSERT_REBUILD (decode and
    rebuild every slot).  The resulting tables are the same."""
    check(lib().gqf_rsqf_set_insert_path(int(path)))


def rsqf_insert_path():
    return lib().gqf_rsqf_get_insert_path()


def sqf_hash():
    return lib().gqf_sqf_get_hash()


def gather_peak(buffer_bytes, n=1 << 24, reps=5):
    """Random 64-byte sector reads per second over a buffer of buffer_bytes
    (gqf_gather_bench): the peak lookups are measured against."""
    ms = C.c_float(0)
    check(lib().gqf_gather_bench(int(buffer_bytes), int(n), int(reps), C.byref(ms)))
    return n / (ms.value / 1e3)


def stream_peak(nbytes=2 << 30, reps=5):
    """Streaming HBM bandwidth in GB/s over buffers of nbytes (gqf_stream_bench):
    read, copy (read + write bytes counted) and write."""
    r, c, w = C.c_float(0), C.c_float(0), C.c_float(0)
    check(lib().gqf_stream_bench(int(nbytes), int(reps), C.byref(r), C.byref(c), C.byref(w)))
    return {"read": nbytes / (r.value / 1e3) / 1e9, "copy": 2 * nbytes / (c.value / 1e3) / 1e9,
            "write": nbytes / (w.value / 1e3) / 1e9, "bytes": int(nbytes), "reps": int(reps)}


def mt19937(seed, n, skip=0):
    """The reference drivers' key stream (init_genrand + genrand_int32)."""
    out = np.empty(int(n), dtype=np.uint32)
    lib().gqf_mt19937_fill(seed, skip, int(n), out.ctypes.data)
    return out


class DeviceArray:
    """A device allocation owned by Python (freed on garbage collection)."""

    def __init__(self, nbytes, dtype=np.uint32, count=None):
        self.dtype = np.dtype(dtype)
        self.nbytes = int(nbytes)
        self.count = int(count if count is not None else self.nbytes // self.dtype.itemsize)
        p = C.c_void_p()
        check(lib().gqf_malloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def empty(cls, count, dtype=np.uint32):
        return cls(int(count) * np.dtype(dtype).itemsize, dtype, count)

    @classmethod
    def from_host(cls, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype)
        d = cls(a.nbytes, a.dtype, a.size)
        if a.nbytes:
            check(lib().gqf_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self, count=None):
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if out.nbytes:
            check(lib().gqf_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def fill(self, byte):
        check(lib().gqf_memset(self.ptr, int(byte), self.nbytes))

    def __len__(self):
        return self.count

    def free(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gqf_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):  # torch tensor on the device
        return x.data_ptr()
    raise TypeError(f"expected a device array, got {type(x)}")


def _n(keys, n):
    if n is not None:
        return int(n)
    return len(keys)


class Stream:
    """A non-blocking HIP stream (default priority).  Calls on different
    streams, from different threads, run concurrently; pass it as `stream=`."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().gqf_stream_create(C.byref(h)))
        self.h = h.value

    @property
    def _as_parameter_(self):
        return C.c_void_p(self.h)

    def destroy(self):
        if self.h is not None:
            check(lib().gqf_stream_destroy(self.h))
            self.h = None


class SQF:
    """Standard quotient filter (quotientFilter.cuh:25-64 API)."""

    def __init__(self, q, r):
        self.h = _SQF()
        check(lib().gqf_sqf_init(C.byref(self.h), q, r))
        self.q, self.r = q, r

    # reference names -------------------------------------------------
    @classmethod
    def initFilterGPU(cls, q, r):
        return cls(q, r)

    @staticmethod
    def calcNumSlotsGPU(q, r):
        """Table size in bytes (the reference function's name)."""
        return lib().gqf_sqf_table_bytes(q, r)

    @staticmethod
    def num_slots(q):
        return lib().gqf_sqf_num_slots(q)

    def _ms(self, fn, *args):
        ms = C.c_float(0)
        check(fn(*args, C.byref(ms)))
        return ms.value

    def bulk_build(self, keys, n=None, NoDuplicates=True, method=BUILD_SEGMENTED_LAYOUTS, stream=None):
        return self._ms(lib().gqf_sqf_bulk_build, self.h, method, _n(keys, n), _ptr(keys), int(bool(NoDuplicates)), stream)

    def bulkBuildSegmentedLayouts(self, keys, n=None, NoDuplicates=True):
        return self.bulk_build(keys, n, NoDuplicates, BUILD_SEGMENTED_LAYOUTS)

    def bulkBuildParallelMerging(self, keys, n=None, NoDuplicates=True):
        return self.bulk_build(keys, n, NoDuplicates, BUILD_PARALLEL_MERGING)

    def bulkBuildSequentialShifts(self, keys, n=None, NoDuplicates=True):
        return self.bulk_build(keys, n, NoDuplicates, BUILD_SEQUENTIAL_SHIFTS)

    bulkBuildEarlyExit = bulkBuildSequentialShifts  # README.md:37 name

    def insert(self, keys, n=None, stream=None):
        return self._ms(lib().gqf_sqf_insert, self.h, _n(keys, n), _ptr(keys), stream)

    def insertViaMerge(self, new_keys, n_new=None, NoDuplicates=True, old_keys=None, n_old=0, stream=None):
        return self._ms(lib().gqf_sqf_insert_via_merge, self.h, _ptr(old_keys), int(n_old), _ptr(new_keys),
                        _n(new_keys, n_new), int(bool(NoDuplicates)), stream)

    def lookup(self, keys, out, n=None, sorted=False, stream=None):
        return self._ms(lib().gqf_sqf_lookup, self.h, int(bool(sorted)), _n(keys, n), _ptr(keys), _ptr(out), stream)

    def launchUnsortedLookups(self, keys, out, n=None):
        return self.lookup(keys, out, n, sorted=False)

    def launchSortedLookups(self, keys, out, n=None):
        return self.lookup(keys, out, n, sorted=True)

    def superclusterDeletes(self, keys, n=None, stream=None):
        return self._ms(lib().gqf_sqf_delete, self.h, _n(keys, n), _ptr(keys), stream)

    delete = superclusterDeletes

    def mergeTwoFilters(self, other, NoDuplicates=True, stream=None):
        return self._ms(lib().gqf_sqf_merge, self.h, other.h, int(bool(NoDuplicates)), stream)

    # host views --------------------------------------------------------
    def table_bytes(self):
        nb = lib().gqf_sqf_table_bytes(self.q, self.r)
        out = np.empty(nb, dtype=np.uint8)
        check(lib().gqf_memcpy_dtoh(out.ctypes.data, self.h.table, nb))
        return out

    def load_table(self, table):
        t = np.ascontiguousarray(table, dtype=np.uint8)
        assert t.size == lib().gqf_sqf_table_bytes(self.q, self.r)
        check(lib().gqf_memcpy_htod(self.h.table, t.ctypes.data, t.nbytes))

    def clear(self):
        check(lib().gqf_memset(self.h.table, 0, lib().gqf_sqf_table_bytes(self.q, self.r)))

    def destroy(self):
        if getattr(self, "h", None) is not None and self.h.table and _lib is not None:
            _lib.gqf_sqf_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class RSQF:
    """Rank-and-select quotient filter (RSQF.cuh:21-67 API, runtime r)."""

    def __init__(self, q, r=5):
        self.h = _RSQF()
        check(lib().gqf_rsqf_init(C.byref(self.h), q, r))
        self.q, self.r = q, r

    @classmethod
    def initCQFGPU(cls, q, r=5):
        return cls(q, r)

    @staticmethod
    def calcNumBlocksGPU(q):
        return lib().gqf_rsqf_num_blocks(q)

    def _ms(self, fn, *args):
        ms = C.c_float(0)
        check(fn(*args, C.byref(ms)))
        return ms.value

    def bulk_build(self, keys, n=None, stream=None):
        return self._ms(lib().gqf_rsqf_bulk_build, self.h, _n(keys, n), _ptr(keys), stream)

    bulkBuild = bulk_build

    def insertGPU(self, keys, out=None, n=None, stream=None):
        return self._ms(lib().gqf_rsqf_insert, self.h, _n(keys, n), _ptr(keys), _ptr(out), stream)

    insert = insertGPU

    def insert_report(self, keys, out=None, n=None, stream=None):
        """insertGPU with the reference's per-item failures: a batch that does
        not fit is inserted in part, out gets -1 for each key left out.
        Returns (ms, number of keys left out)."""
        failed = C.c_uint32(0)
        ms = C.c_float(0)
        check(lib().gqf_rsqf_insert_report(self.h, _n(keys, n), _ptr(keys), _ptr(out), C.byref(failed), stream,
                                           C.byref(ms)))
        return ms.value, failed.value

    def lookup(self, keys, out, n=None, sorted=True, stream=None):
        return self._ms(lib().gqf_rsqf_lookup, self.h, int(bool(sorted)), _n(keys, n), _ptr(keys), _ptr(out), stream)

    def launchLookups(self, keys, out, n=None):
        return self.lookup(keys, out, n, sorted=True)

    def launchUnsortedLookups(self, keys, out, n=None):
        return self.lookup(keys, out, n, sorted=False)

    def count(self, keys, out, n=None, stream=None):
        """Stored copies of each key's fingerprint (gqf_rsqf_count)."""
        return self._ms(lib().gqf_rsqf_count, self.h, _n(keys, n), _ptr(keys), _ptr(out), stream)

    def deleteGPU(self, keys, n=None, stream=None):
        return self._ms(lib().gqf_rsqf_delete, self.h, _n(keys, n), _ptr(keys), stream)

    delete = deleteGPU

    def dump(self):
        """Reference packed-block layout (RSQF.cuh:24-30) as host bytes."""
        out = np.empty(lib().gqf_rsqf_ref_bytes(self.q, self.r), dtype=np.uint8)
        check(lib().gqf_rsqf_dump(self.h, out.ctypes.data))
        return out

    def load(self, ref_layout):
        t = np.ascontiguousarray(ref_layout, dtype=np.uint8)
        assert t.size == lib().gqf_rsqf_ref_bytes(self.q, self.r)
        check(lib().gqf_rsqf_load(self.h, t.ctypes.data))

    def clear(self):
        check(lib().gqf_memset(self.h.blocks, 0, lib().gqf_rsqf_table_bytes(self.q, self.r)))

    def destroy(self):
        if getattr(self, "h", None) is not None and self.h.blocks and _lib is not None:
            _lib.gqf_rsqf_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def serve(filters, requests, coalesce=True, hbm_budget=64 << 30, sorted_lookups=False, max_batch=0):
    """Run a request stream over resident RSQFs (gqf_rsqf_serve).

    ``requests`` is a ctypes array of :class:`Request` (filled in place with
    start/done times, status and call ids).  Returns the :class:`ServeStats`."""
    arr = (_RSQF * len(filters))(*[f.h for f in filters])
    cfg = ServeConfig(int(hbm_budget), int(bool(coalesce)), int(bool(sorted_lookups)), int(max_batch))
    st = ServeStats()
    check(lib().gqf_rsqf_serve(arr, len(filters), requests, len(requests), C.byref(cfg), C.byref(st)))
    return st


class SQFBatch:
    """Many SQFs per launch (gqf_sqf_batch_*).

    ``filters``: initialised :class:`SQF` objects of one r.  ``keys`` /
    ``key_offsets``: every filter's build keys back to back in one
    :class:`DeviceArray` (offsets: F + 1 host integers); ``probe`` /
    ``probe_offsets`` / ``out`` likewise for lookups.  Creating the batch
    autotunes and captures the pipelines (and leaves every table built)."""

    def __init__(self, filters, keys=None, key_offsets=None, probe=None, probe_offsets=None, out=None,
                 no_dup=False, stream=None):
        self.filters = list(filters)
        F = len(self.filters)
        self._arr = (_SQF * F)(*[f.h for f in self.filters])
        self._koff = (C.c_uint64 * (F + 1))(*[int(x) for x in key_offsets]) if key_offsets is not None else None
        self._poff = (C.c_uint64 * (F + 1))(*[int(x) for x in probe_offsets]) if probe_offsets is not None else None
        self._hold = (keys, probe, out)
        self.h = _vp()
        check(lib().gqf_sqf_batch_create(C.byref(self.h), self._arr, F, int(bool(no_dup)), _ptr(keys), self._koff,
                                         _ptr(probe), self._poff, _ptr(out), stream))

    def build(self, use_graph=True):
        """Returns (ms, per-filter statuses)."""
        st = (C.c_int * len(self.filters))()
        ms = C.c_float(0)
        check(lib().gqf_sqf_batch_build(self.h, int(bool(use_graph)), st, C.byref(ms)))
        return ms.value, list(st)

    def lookup(self, use_graph=True):
        ms = C.c_float(0)
        check(lib().gqf_sqf_batch_lookup(self.h, int(bool(use_graph)), C.byref(ms)))
        return ms.value

    def lookup_walk(self):
        """The same lookups with every walk run to its end (no window index):
        the per-filter algorithm in one launch."""
        ms = C.c_float(0)
        check(lib().gqf_sqf_batch_lookup_walk(self.h, C.byref(ms)))
        return ms.value

    def info(self):
        i = BatchInfo()
        check(lib().gqf_sqf_batch_info(self.h, C.byref(i)))
        return i.as_dict()

    def destroy(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.gqf_sqf_batch_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def image_info(image):
    """(layout, q, r, hash) of a table image (gqf_image_info)."""
    b = np.frombuffer(bytes(image), dtype=np.uint8)
    lay, hs = C.c_int(), C.c_int()
    q, r = C.c_uint(), C.c_uint()
    check(lib().gqf_image_info(b.ctypes.data, b.size, C.byref(lay), C.byref(q), C.byref(r), C.byref(hs)))
    return lay.value, q.value, r.value, hs.value


def image_save(f, layout=None):
    """Versioned image of an SQF (layout IMAGE_SQF_GPU or IMAGE_SQF_CPU) or RSQF table, as bytes."""
    if isinstance(f, RSQF):
        out = np.empty(lib().gqf_image_bytes(IMAGE_RSQF, f.q, f.r), np.uint8)
        check(lib().gqf_rsqf_image_save(f.h, out.ctypes.data, out.size))
    else:
        layout = IMAGE_SQF_GPU if layout is None else layout
        out = np.empty(lib().gqf_image_bytes(layout, f.q, f.r), np.uint8)
        check(lib().gqf_sqf_image_save(f.h, layout, out.ctypes.data, out.size))
    return out.tobytes()


def image_load(f, image):
    b = np.frombuffer(bytes(image), dtype=np.uint8)
    if isinstance(f, RSQF):
        check(lib().gqf_rsqf_image_load(f.h, b.ctypes.data, b.size))
    else:
        check(lib().gqf_sqf_image_load(f.h, b.ctypes.data, b.size))
