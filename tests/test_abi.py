"""CPU-side checks of the C ABI: libgqf.so loads and exports every entry
point include/gqf.h declares; geometry helpers agree with the oracle.  No
compute call needs a GPU here."""
import os
import re

import pytest

from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "gqf.h")).read()
    return sorted(set(re.findall(r"\b(gqf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = _declared()
    for must in ["gqf_sqf_init", "gqf_sqf_bulk_build", "gqf_sqf_insert", "gqf_sqf_insert_via_merge",
                 "gqf_sqf_lookup", "gqf_sqf_delete", "gqf_sqf_merge", "gqf_rsqf_init", "gqf_rsqf_bulk_build",
                 "gqf_rsqf_insert", "gqf_rsqf_lookup", "gqf_rsqf_delete", "gqf_rsqf_dump", "gqf_rsqf_load"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import gpuquotientfilters_amd as G

    L = G.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # and the binding's table covers the header
    assert set(_declared()) <= set(G.EXPORTED)


def test_geometry_matches_oracle(orc):
    import gpuquotientfilters_amd as G

    L = G.lib()
    for q in (8, 12, 18, 20, 26):
        for r in (5, 13):
            if q + r <= 32:
                assert L.gqf_sqf_table_bytes(q, r) == orc.sqf_table_bytes(q, r)
        assert L.gqf_sqf_num_slots(q) == orc.sqf_num_slots(q)
        assert L.gqf_rsqf_num_blocks(q) == orc.rsqf_num_blocks(q)
        for r in (5, 8):
            assert L.gqf_rsqf_ref_bytes(q, r) == orc.rsqf_table_bytes(q, r)


def test_version_and_errors():
    import gpuquotientfilters_amd as G

    assert G.lib().gqf_version().decode().startswith("gqf")
    assert G.lib().gqf_status_string(2).decode() == "capacity exceeded"


def test_invalid_geometry_rejected_without_gpu():
    import ctypes as C

    import gpuquotientfilters_amd as G

    h = G._SQF()
    # packed widths reach 16-bit elements; wider ones, q + r > 32 and q > 28
    # are rejected before any device call
    assert G.lib().gqf_sqf_init(C.byref(h), 10, 14) == G.ERR_INVALID
    assert G.lib().gqf_sqf_init(C.byref(h), 25, 8) == G.ERR_INVALID
    assert G.lib().gqf_sqf_init(C.byref(h), 29, 1) == G.ERR_INVALID
    assert G.lib().gqf_sqf_init(C.byref(h), 10, 0) == G.ERR_INVALID
    # sizes of packed tables: S (r+3)-bit slots in 32-bit words, plus one
    for q, r in [(10, 6), (12, 8), (20, 9), (26, 4)]:
        S = G.lib().gqf_sqf_num_slots(q)
        assert G.lib().gqf_sqf_table_bytes(q, r) == 4 * ((S * (r + 3) + 31) // 32 + 1)
    h2 = G._RSQF()
    assert G.lib().gqf_rsqf_init(C.byref(h2), 10, 40) == G.ERR_INVALID


def test_sqf_hash_selector():
    import gpuquotientfilters_amd as G
    assert G.sqf_hash() == G.HASH_AP
    G.set_sqf_hash(G.HASH_FNV)
    assert G.sqf_hash() == G.HASH_FNV
    G.set_sqf_hash(G.HASH_AP)
    with __import__("pytest").raises(G.GQFError):
        G.set_sqf_hash(7)
    assert G.sqf_hash() == G.HASH_AP


def test_library_key_stream_is_the_reference_stream():
    """gqf_mt19937_fill generates every bench input: it must reproduce the
    reference drivers' init_genrand + genrand_int32 stream (golden vectors
    written by the compiled mt19937ar.cpp), also after skipping outputs."""
    import numpy as np

    import gpuquotientfilters_amd as G
    from conftest import golden

    g = golden("mt19937.npz")
    assert (G.mt19937(1, 4096) == g["seed1"]).all()
    assert (G.mt19937(7, 1024) == g["seed7"]).all()
    assert (G.mt19937(1, 1000, skip=3000) == g["seed1"][3000:4000]).all()
    assert (G.mt19937(7, 1, skip=1023) == g["seed7"][1023:]).all()
    # C++ standard known answer: the 10000th output of mt19937 from seed 5489
    assert int(G.mt19937(5489, 1, skip=9999)[0]) == 4123659995
    assert G.mt19937(3, 0).shape == (0,)
    assert np.array_equal(G.mt19937(3, 5), G.mt19937(3, 7)[:5])
