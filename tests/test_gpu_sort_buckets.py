"""Batches of 4097 .. 2^20 keys (2^19 when fingerprints are wider than 32
bits) are sorted by one scatter on the top 8 fingerprint bits followed by one
workgroup per bucket.  These tests drive that form through the operations
that sort (builds, inserts, sorted lookups), bit-exact against the oracle, at
its size bounds.  A bucket holds 8192 32-bit or 4096 64-bit keys; a key
repeated more often than that makes the workgroup sort its bucket by LSD
passes over global memory; with 13, 17, 19, 20 or 26 bits below the top
digit those are 2, 3, 3, 3 or 4 passes, so both the even count (a final copy)
and the odd one are run."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from gpu_helpers import dev, fresh_keys, require_gpu, rsqf_lookup, say, sqf_canon, sqf_lookup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def _with_hot(seed, n, hot_copies):
    keys = O.mt19937(seed, n)
    if hot_copies:
        keys[:hot_copies] = keys[n - 1]  # one key n - hot_copies + 1 times
        keys = keys[np.random.default_rng(seed).permutation(n)]
    return keys


@pytest.mark.parametrize("q,r,n,hot", [
    (13, 5, 4097, 0),             # just past the one-workgroup sort; 10 bits below the top digit
    (16, 5, 30000, 0),
    (16, 5, 30000, 9000),         # bucket overflow, 2 passes (ends with a copy)
    (20, 5, 1 << 19, 0),
    (20, 5, 600000, 20000),       # bucket overflow, 3 passes
    (21, 5, 1 << 20, 0),          # largest bucket-sorted batch
    (21, 5, (1 << 20) + 1, 0),    # smallest batch sorted by LSD passes over the whole batch
    (12, 13, 3000, 0),            # one-workgroup sort
    (14, 13, 14000, 9000),        # 19 bits below the top digit: 3 passes
])
@pytest.mark.parametrize("no_dup", [True, False])
def test_sqf_build_bucket_sizes(q, r, n, hot, no_dup):
    say("sqf bucket build", q, r, n, hot, no_dup)
    keys = _with_hot(41, n, hot)
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=no_dup)
    want = sqf_canon(keys, q, r, no_dup)
    got = f.table_bytes()
    assert (got == want).all(), np.flatnonzero(got != want)[:8]


def test_sqf_insert_batches_with_hot_keys():
    # inserts sort each batch (and keep one copy of each new fingerprint);
    # every batch carries up to 10000 copies of one key
    q, r = 20, 5
    base = O.mt19937(43, 200000)
    f = G.SQF(q, r)
    f.bulk_build(dev(base), NoDuplicates=True)
    all_keys = [base]
    for i, n in enumerate([5000, 70000, 600000]):
        b = _with_hot(44 + i, n, min(10000, n // 2))
        f.insert(dev(b))
        all_keys.append(b)
    want = sqf_canon(np.concatenate(all_keys), q, r, True)
    assert (f.table_bytes() == want).all()


@pytest.mark.parametrize("n_probe", [300000, 700000])
def test_sqf_sorted_lookups_with_repeated_queries(n_probe):
    # the sorted path sorts fingerprints with their query indices
    q, r = 20, 5
    keys = O.mt19937(45, O.key_count(0.6, q))
    table = sqf_canon(keys, q, r, True)
    f = G.SQF(q, r)
    f.load_table(table)
    miss = fresh_keys(46, 1, keys)[0]
    rest = n_probe - 60000
    probe = np.concatenate([keys[: rest // 2], fresh_keys(47, rest - rest // 2, keys),
                            np.full(30000, keys[7], np.uint32), np.full(30000, miss, np.uint32)])
    probe = probe[np.random.default_rng(48).permutation(len(probe))]
    want = O.sqf_lookup_keys(table, q, r, probe)
    assert (sqf_lookup(f, probe, sorted=True) == want).all()
    assert (sqf_lookup(f, probe, sorted=False) == want).all()


def _rsqf_canon(keys, q, r):
    st, t = O.rsqf_canon(O.fingerprints(keys, q, r), q, r)
    assert st == O.OK
    return t


@pytest.mark.parametrize("q,r", [(20, 8), (24, 10)])
def test_rsqf_build_insert_with_hot_keys(q, r):
    """20 bits (32-bit sort keys) and 26 bits (64-bit) below the top digit;
    9000 copies of one key overflow both bucket sizes, 5000 the 4096-key
    one."""
    say("rsqf bucket build/insert", q, r)
    base = _with_hot(51, 300000, 9000)
    f = G.RSQF(q, r)
    f.bulk_build(dev(base))
    extra = _with_hot(52, 60000, 5000)
    f.insertGPU(dev(extra))
    assert (f.dump() == _rsqf_canon(np.concatenate([base, extra]), q, r)).all()


@pytest.mark.parametrize("q,r", [(20, 8), (24, 10)])
def test_rsqf_sorted_lookups_with_repeated_queries(q, r):
    say("rsqf bucket sorted lookups", q, r)
    keys = O.mt19937(53, 300000)
    t = _rsqf_canon(keys, q, r)
    f = G.RSQF(q, r)
    f.load(t)
    miss = fresh_keys(54, 1, keys)[0]
    probe = np.concatenate([keys[:100000], fresh_keys(55, 150000, keys),
                            np.full(20000, keys[3], np.uint32), np.full(20000, miss, np.uint32)])
    probe = probe[np.random.default_rng(56).permutation(len(probe))]
    want = O.rsqf_lookup_keys(t, q, r, probe)
    assert (rsqf_lookup(f, probe, sorted=True) == want).all()
    assert (rsqf_lookup(f, probe, sorted=False) == want).all()
