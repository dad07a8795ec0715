"""CPU oracle for the quotient-filter engine -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker / CPU baseline.  The
product library (``gpuquotientfilters_amd``) never imports it.

``liboracle.so`` is the project's C++ restatement of the reference CPU_version
algorithms (``qf_oracle.cpp``, each function citing the reference file:line it
follows).  ``_ref/`` holds the reference CPU sources compiled by ``Makefile``
(``make ref``) when the reference tree is present; ``RefSQF``/``RefRSQF`` wrap
them for fixture generation and cross-checks.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NOT_FOUND = 0xFFFFFFFF
OK, ERR_INVALID, ERR_CAPACITY = 0, 1, 2

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_sz, _u, _i, _d = C.c_size_t, C.c_uint, C.c_int, C.c_double

_SIGS = {
    "orc_mt19937": (None, [C.c_uint32, C.c_uint64, _sz, _u32p]),
    "orc_ap_rounds": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "orc_fmix64": (C.c_uint64, [C.c_uint64]),
    "orc_fp_ap": (C.c_uint64, [C.c_uint32, _u, _u]),
    "orc_fp_fnv": (C.c_uint64, [C.c_uint32, _u, _u]),
    "orc_fingerprints": (None, [_u32p, _sz, _u, _u, _i, _u64p]),
    "orc_sqf_num_slots": (C.c_uint32, [_u]),
    "orc_sqf_table_bytes": (_sz, [_u, _u]),
    "orc_sqf_cpu_words": (_sz, [_u, _u]),
    "orc_sqf_decode_bytes": (None, [_u8p, _u, C.c_uint32, _u32p]),
    "orc_sqf_decode_words": (None, [_u32p, _sz, _u, C.c_uint32, _u32p]),
    "orc_sqf_canon": (_i, [_u64p, _sz, _u, _u, _i, _u8p]),
    "orc_sqf_lookup_fp": (C.c_uint32, [_u8p, _u, _u, C.c_uint64]),
    "orc_sqf_insert_fp": (C.c_uint32, [_u8p, _u, _u, C.c_uint64, C.POINTER(C.c_int)]),
    "orc_sqf_delete_fp": (_i, [_u8p, _u, _u, C.c_uint64]),
    "orc_sqf_extract": (_sz, [_u8p, _u, _u, _u64p]),
    "orc_sqf_lookup_keys": (None, [_u8p, _u, _u, _i, _u32p, _sz, _u32p]),
    "orc_sqf_insert_keys": (_i, [_u8p, _u, _u, _i, _u32p, _sz]),
    "orc_sqf_delete_keys": (_sz, [_u8p, _u, _u, _i, _u32p, _sz]),
    "orc_rsqf_num_blocks": (C.c_uint32, [_u]),
    "orc_rsqf_block_bytes": (_sz, [_u]),
    "orc_rsqf_table_bytes": (_sz, [_u, _u]),
    "orc_rank": (C.c_uint32, [C.c_uint64, _u]),
    "orc_select": (C.c_uint32, [C.c_uint64, _u]),
    "orc_rsqf_get_rem": (C.c_uint32, [_u8p, _u, C.c_uint32, _u]),
    "orc_rsqf_canon": (_i, [_u64p, _sz, _u, _u, _u8p]),
    "orc_rsqf_lookup_fp": (C.c_uint32, [_u8p, _u, _u, C.c_uint64]),
    "orc_rsqf_insert_fp": (C.c_uint32, [_u8p, _u, _u, C.c_uint64]),
    "orc_rsqf_lookup_keys": (None, [_u8p, _u, _u, _u32p, _sz, _u32p]),
    "orc_rsqf_insert_keys": (_sz, [_u8p, _u, _u, _u32p, _sz, _u32p]),
    "orc_rsqf_extract": (_sz, [_u8p, _u, _u, _u64p]),
    "orc_time_sqf_insert": (_d, [_u8p, _u, _u, _u32p, _sz]),
    "orc_time_sqf_lookup": (_d, [_u8p, _u, _u, _u32p, _sz, _u32p, _i]),
    "orc_time_sqf_delete": (_d, [_u8p, _u, _u, _u32p, _sz]),
    "orc_time_rsqf_insert": (_d, [_u8p, _u, _u, _u32p, _sz]),
    "orc_time_rsqf_lookup": (_d, [_u8p, _u, _u, _u32p, _sz, _u32p, _i]),
}


def build(force=False):
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    cmd = ["make", "-s", "-C", HERE]
    if force:
        cmd.append("-B")
    subprocess.run(cmd + ["all"], check=True)
    if os.path.isdir(os.environ.get("GQF_REFERENCE", "/root/reference")):
        subprocess.run(cmd + ["ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        # GQF_ORACLE_LIB: another build of the same source (the UBSan test)
        path = os.environ.get("GQF_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---------------------------------------------------------------- keys / hashes
def mt19937(seed, n, skip=0):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt19937(seed, skip, n, out)
    return out


def fingerprints(keys, q, r, hash_kind=0):
    keys = _u32(keys)
    out = np.empty(len(keys), dtype=np.uint64)
    lib().orc_fingerprints(keys, len(keys), q, r, hash_kind, out)
    return out


def fp_ap(key, q, r):
    return lib().orc_fp_ap(key, q, r)


def key_count(alpha, q):
    """Reference driver rule n = (int)(float alpha * 2^q) (e.g. bulkBuildSegmentedLayoutPerf.cu:51)."""
    return int(np.float32(alpha) * np.float32(1 << q))


# ---------------------------------------------------------------- SQF
def sqf_num_slots(q):
    return lib().orc_sqf_num_slots(q)


def sqf_table_bytes(q, r):
    return lib().orc_sqf_table_bytes(q, r)


def sqf_decode(table, q, r):
    s = sqf_num_slots(q)
    out = np.empty(s, dtype=np.uint32)
    lib().orc_sqf_decode_bytes(np.ascontiguousarray(table, dtype=np.uint8), r, s, out)
    return out


def sqf_decode_words(words, q, r, nslots=None):
    s = sqf_num_slots(q) if nslots is None else nslots
    out = np.empty(s, dtype=np.uint32)
    w = _u32(words)
    lib().orc_sqf_decode_words(w, len(w), r, s, out)
    return out


def sqf_packed_words(q, r):
    """Words of a packed-width GPU table: S (r+3)-bit slots, plus one word."""
    return (sqf_num_slots(q) * (r + 3) + 31) // 32 + 1


def sqf_pack(elems, q, r):
    """Elements (one per slot) -> (r+3)-bit fields packed MSB-first into uint32
    words, the CPU_version's bit order (getElement, quotientFilter.cpp:114-130):
    slot i at bits [i(r+3), (i+1)(r+3)) from the top of word 0.  Inverse of
    sqf_decode_words (checked against it and the reference's own words)."""
    w = r + 3
    e = np.ascontiguousarray(elems, dtype=np.uint32)
    bits = ((e[:, None] >> np.arange(w - 1, -1, -1, dtype=np.uint32)) & 1).astype(np.uint8).ravel()
    nwords = sqf_packed_words(q, r)
    buf = np.zeros(nwords * 32, np.uint8)
    buf[: len(bits)] = bits
    return np.packbits(buf, bitorder="big").view(">u4").astype(np.uint32)


def sqf_canon(fps, q, r, no_dup):
    fps = _u64(fps)
    t = np.zeros(sqf_table_bytes(q, r), dtype=np.uint8)
    st = lib().orc_sqf_canon(fps, len(fps), q, r, int(bool(no_dup)), t)
    return st, t


def sqf_lookup_keys(table, q, r, keys, hash_kind=0):
    keys = _u32(keys)
    out = np.empty(len(keys), dtype=np.uint32)
    lib().orc_sqf_lookup_keys(table, q, r, hash_kind, keys, len(keys), out)
    return out


def sqf_insert_keys(table, q, r, keys, hash_kind=0):
    keys = _u32(keys)
    return lib().orc_sqf_insert_keys(table, q, r, hash_kind, keys, len(keys))


def sqf_delete_keys(table, q, r, keys, hash_kind=0):
    keys = _u32(keys)
    return lib().orc_sqf_delete_keys(table, q, r, hash_kind, keys, len(keys))


def sqf_extract(table, q, r):
    out = np.empty(sqf_num_slots(q), dtype=np.uint64)
    k = lib().orc_sqf_extract(table, q, r, out)
    return out[:k].copy()


# ---------------------------------------------------------------- RSQF
def rsqf_num_blocks(q):
    return lib().orc_rsqf_num_blocks(q)


def rsqf_table_bytes(q, r):
    return lib().orc_rsqf_table_bytes(q, r)


def rsqf_canon(fps, q, r):
    fps = _u64(fps)
    t = np.zeros(rsqf_table_bytes(q, r), dtype=np.uint8)
    st = lib().orc_rsqf_canon(fps, len(fps), q, r, t)
    return st, t


def rsqf_lookup_keys(table, q, r, keys):
    keys = _u32(keys)
    out = np.empty(len(keys), dtype=np.uint32)
    lib().orc_rsqf_lookup_keys(table, q, r, keys, len(keys), out)
    return out


def rsqf_insert_keys(table, q, r, keys):
    keys = _u32(keys)
    out = np.empty(len(keys), dtype=np.uint32)
    fails = lib().orc_rsqf_insert_keys(table, q, r, keys, len(keys), out)
    return fails, out


def rsqf_extract(table, q, r):
    out = np.empty(rsqf_num_blocks(q) * 64, dtype=np.uint64)
    k = lib().orc_rsqf_extract(table, q, r, out)
    return out[:k].copy()


def rsqf_fields(table, q, r):
    """Split a reference-layout packed RSQF table into (offsets, occupieds, runEnds, remainder bytes)."""
    nb = rsqf_num_blocks(q)
    bb = lib().orc_rsqf_block_bytes(r)
    t = np.ascontiguousarray(table, dtype=np.uint8).reshape(nb, bb)
    off = t[:, 0].copy()
    occ = t[:, 1:9].copy().view("<u8").reshape(nb)
    ends = t[:, 9:17].copy().view("<u8").reshape(nb)
    rem = t[:, 17:].copy()
    return off, occ, ends, rem


# ---------------------------------------------------------------- reference CPU build (oracle/_ref)
def _ref_lib(name):
    path = os.path.join(HERE, "_ref", name)
    return C.CDLL(path) if os.path.exists(path) else None


class RefSQF:
    """The reference CPU_version SQF (FNV hash, packed uint32 words), compiled from its sources."""

    _L = None

    @classmethod
    def available(cls):
        return cls.lib() is not None

    @classmethod
    def lib(cls):
        if cls._L is None:
            L = _ref_lib("libqfref_sqf.so")
            if L is None:
                return None
            L.refsqf_new.restype = C.c_void_p
            L.refsqf_new.argtypes = [_u, _u]
            L.refsqf_free.argtypes = [C.c_void_p]
            L.refsqf_words.restype = _sz
            L.refsqf_words.argtypes = [_u, _u]
            L.refsqf_table.restype = C.POINTER(C.c_uint32)
            L.refsqf_table.argtypes = [C.c_void_p]
            L.refsqf_fnv.restype = C.c_uint32
            L.refsqf_fnv.argtypes = [C.c_uint32, C.c_uint32]
            L.refsqf_insert_many.argtypes = [C.c_void_p, _u32p, _sz, _u32p]
            L.refsqf_lookup_many.argtypes = [C.c_void_p, _u32p, _sz, _i32p]
            L.refmt_fill.argtypes = [C.c_ulong, _sz, _u32p]
            cls._L = L
        return cls._L

    def __init__(self, q, r):
        self.q, self.r = q, r
        self.h = self.lib().refsqf_new(q, r)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().refsqf_free(self.h)
            self.h = None

    def words(self):
        n = self.lib().refsqf_words(self.q, self.r)
        p = self.lib().refsqf_table(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def insert(self, keys):
        keys = _u32(keys)
        out = np.empty(len(keys), dtype=np.uint32)
        self.lib().refsqf_insert_many(self.h, keys, len(keys), out)
        return out

    def lookup(self, keys):
        keys = _u32(keys)
        out = np.empty(len(keys), dtype=np.int32)
        self.lib().refsqf_lookup_many(self.h, keys, len(keys), out)
        return out

    @classmethod
    def fnv(cls, key, table_size):
        return cls.lib().refsqf_fnv(key, table_size)

    @classmethod
    def mt(cls, seed, n):
        out = np.empty(n, dtype=np.uint32)
        cls.lib().refmt_fill(seed, n, out)
        return out


class RefRSQF:
    """The reference CPU_version RSQF (AP hash, RBITS from its header), compiled from its sources."""

    _L = None

    @classmethod
    def available(cls):
        return cls.lib() is not None

    @classmethod
    def lib(cls):
        if cls._L is None:
            L = _ref_lib("libqfref_rsqf.so")
            if L is None:
                return None
            L.refrsqf_rbits.restype = _u
            L.refrsqf_block_bytes.restype = _sz
            L.refrsqf_new.restype = C.c_void_p
            L.refrsqf_new.argtypes = [_u]
            L.refrsqf_free.argtypes = [C.c_void_p]
            L.refrsqf_num_blocks.restype = _sz
            L.refrsqf_num_blocks.argtypes = [_u]
            L.refrsqf_blocks.restype = C.POINTER(C.c_uint8)
            L.refrsqf_blocks.argtypes = [C.c_void_p]
            L.refrsqf_ap.restype = C.c_uint32
            L.refrsqf_ap.argtypes = [C.c_uint32, C.c_uint32]
            L.refrsqf_rank.restype = C.c_uint32
            L.refrsqf_rank.argtypes = [C.c_uint64, _u]
            L.refrsqf_select.restype = C.c_uint32
            L.refrsqf_select.argtypes = [C.c_uint64, _u]
            L.refrsqf_insert_many.argtypes = [C.c_void_p, _u32p, _sz, _i32p]
            L.refrsqf_lookup_many.argtypes = [C.c_void_p, _u32p, _sz, _i32p]
            cls._L = L
        return cls._L

    def __init__(self, q):
        self.q = q
        self.r = self.lib().refrsqf_rbits()
        self.h = self.lib().refrsqf_new(q)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().refrsqf_free(self.h)
            self.h = None

    def table(self):
        n = self.lib().refrsqf_num_blocks(self.q) * self.lib().refrsqf_block_bytes()
        p = self.lib().refrsqf_blocks(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def insert(self, keys):
        keys = _u32(keys)
        out = np.empty(len(keys), dtype=np.int32)
        self.lib().refrsqf_insert_many(self.h, keys, len(keys), out)
        return out

    def lookup(self, keys):
        keys = _u32(keys)
        out = np.empty(len(keys), dtype=np.int32)
        self.lib().refrsqf_lookup_many(self.h, keys, len(keys), out)
        return out
