"""gqf_sqf_bulk_build_noalloc: the SQF build over the library's fixed device
scratch array (no allocation), called through the C ABI and checked against
gqf_sqf_bulk_build (the same keys into a second filter: identical tables and
statuses) and the oracle's canonical layout -- two calls in a row, the
second with fewer keys, so leftovers of the first call in the shared array
must not leak into the second."""
import ctypes as C

import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from gpu_helpers import dev, require_gpu, say, sqf_canon

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def noalloc(f, d_keys, n, no_dup):
    ms = C.c_float(0)
    return G.lib().gqf_sqf_bulk_build_noalloc(f.h, 0, int(n), d_keys.ptr if n else None, int(no_dup), None,
                                             C.byref(ms))


def plain(f, d_keys, n, no_dup):
    return G.lib().gqf_sqf_bulk_build(f.h, 0, int(n), d_keys.ptr if n else None, int(no_dup), None, None)


@pytest.mark.parametrize("q,r,no_dup", [(21, 5, True), (21, 5, False), (16, 13, True)])
def test_two_calls_match_the_allocating_build(q, r, no_dup):
    n1 = min(G.NOALLOC_MAX_KEYS, O.key_count(0.8, q))
    keys = O.mt19937(41, n1)
    keys[: n1 // 10] = keys[n1 // 10: 2 * (n1 // 10)]  # duplicates for the Dup / NoDup rules
    small = O.mt19937(42, 3000)
    a, b = G.SQF(q, r), G.SQF(q, r)
    for step, k in enumerate((keys, small)):
        say("noalloc build", q, r, no_dup, "call", step, "n", len(k))
        d = dev(k)
        assert noalloc(a, d, len(k), no_dup) == G.OK, G.lib().gqf_last_error()
        assert plain(b, d, len(k), no_dup) == G.OK
        ta = a.table_bytes()
        assert (ta == b.table_bytes()).all()
        assert (ta == sqf_canon(k, q, r, no_dup)).all()


def test_capacity_error_and_empty_input():
    q, r = 10, 5
    f, g = G.SQF(q, r), G.SQF(q, r)
    ok = O.mt19937(43, 500)
    d = dev(ok)
    assert noalloc(f, d, len(ok), True) == G.OK
    assert plain(g, d, len(ok), True) == G.OK
    before = f.table_bytes()
    big = dev(O.mt19937(44, 5000))  # more keys than the 1126 slots
    say("noalloc capacity", q, r)
    assert noalloc(f, big, 5000, False) == G.ERR_CAPACITY
    assert plain(g, big, 5000, False) == G.ERR_CAPACITY
    assert (f.table_bytes() == before).all() and (g.table_bytes() == before).all()
    # an empty build clears the table, as the allocating build does
    assert noalloc(f, d, 0, True) == G.OK and plain(g, d, 0, True) == G.OK
    assert (f.table_bytes() == 0).all() and (g.table_bytes() == 0).all()


def test_too_many_keys_is_refused_before_any_device_work():
    f = G.SQF(22, 5)
    d = G.DeviceArray.empty(G.NOALLOC_MAX_KEYS + 1)
    assert noalloc(f, d, G.NOALLOC_MAX_KEYS + 1, True) == G.ERR_INVALID
    assert "at most" in G.lib().gqf_last_error().decode()
