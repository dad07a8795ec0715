"""Sorts of skewed and crowded batches, seen through the operations that
sort (builds, inserts, sorted lookups), bit-exact against the oracle:

  * sizes at the bounds of the sort's forms (4096 keys and fewer: one
    workgroup; 4097 to 2^20 keys: a partition on the top 8 bits, then one
    workgroup per segment, or the chunked LSD passes when a segment holds
    more than 8192 keys; above 2^20: the chunked LSD passes) and around
    2^18;
  * one key repeated thousands of times (a digit value holding most of a
    chunk);
  * thousands of distinct keys whose fingerprints share their top 8 bits,
    placed so that evenly spaced samples miss them;
  * Zipf ranks as keys (small integers): the reference hash crowds their
    fingerprints into a few top-bit values (one value held 72% of a 2^18
    batch), which any sort that splits by top bits has to survive.
"""
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
    (13, 5, 4097, 0),             # just past the one-workgroup sort
    (16, 5, 30000, 0),
    (16, 5, 30000, 9000),         # 9001 copies of one key
    (20, 5, 1 << 18, 0),
    (20, 5, (1 << 18) + 1, 0),
    (20, 5, 250000, 20000),
    (12, 13, 3000, 0),            # one-workgroup sort
    (14, 13, 14000, 9000),        # 16-bit slots
])
@pytest.mark.parametrize("no_dup", [True, False])
def test_sqf_build_sizes_and_hot_keys(q, r, n, hot, no_dup, fresh_stream):
    say("sqf sort build", q, r, n, hot, no_dup)
    keys = _with_hot(41, n, hot)
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=no_dup, stream=fresh_stream)
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
    for i, n in enumerate([5000, 70000, 250000]):
        b = _with_hot(44 + i, n, min(10000, n // 2))
        f.insert(dev(b))
        all_keys.append(b)
    want = sqf_canon(np.concatenate(all_keys), q, r, True)
    assert (f.table_bytes() == want).all()


@pytest.mark.parametrize("n_probe", [100000, 260000])
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
    """32-bit and 64-bit sort keys; 9000 and 5000 copies of one key."""
    say("rsqf sort build/insert", q, r)
    base = _with_hot(51, 250000, 9000)
    f = G.RSQF(q, r)
    f.bulk_build(dev(base))
    extra = _with_hot(52, 60000, 5000)
    f.insertGPU(dev(extra))
    assert (f.dump() == _rsqf_canon(np.concatenate([base, extra]), q, r)).all()


@pytest.mark.parametrize("q,r", [(20, 8), (24, 10)])
def test_rsqf_sorted_lookups_with_repeated_queries(q, r):
    say("rsqf sorted lookups, repeated", q, r)
    keys = O.mt19937(53, 300000)
    t = _rsqf_canon(keys, q, r)
    f = G.RSQF(q, r)
    f.load(t)
    miss = fresh_keys(54, 1, keys)[0]
    probe = np.concatenate([keys[:80000], fresh_keys(55, 140000, keys),
                            np.full(20000, keys[3], np.uint32), np.full(20000, miss, np.uint32)])
    probe = probe[np.random.default_rng(56).permutation(len(probe))]
    want = O.rsqf_lookup_keys(t, q, r, probe)
    assert (rsqf_lookup(f, probe, sorted=True) == want).all()
    assert (rsqf_lookup(f, probe, sorted=False) == want).all()


def _crowded(seed, q, r, n_crowd, n_rest):
    """n_crowd distinct keys of the lowest fingerprints (top 8 bits zero) and
    n_rest others, laid out so that the 1024 evenly spaced positions
    (i * n // 1024) all hold one of the others."""
    bits = q + r
    pool = O.mt19937(seed, 600 * n_crowd)
    fp = O.fingerprints(pool, q, r)
    low = np.unique(pool[(fp >> np.uint64(bits - 8)) == 0])
    assert len(low) >= n_crowd
    low = low[np.random.default_rng(seed).permutation(len(low))[:n_crowd]]
    other = O.mt19937(seed + 1, 2 * n_rest)
    other = other[(O.fingerprints(other, q, r) >> np.uint64(bits - 8)) != 0][:n_rest]
    n = n_crowd + n_rest
    sampled = np.unique(np.arange(1024, dtype=np.int64) * n // 1024)
    assert len(other) == n_rest >= len(sampled)
    keys = np.empty(n, np.uint32)
    is_s = np.zeros(n, bool)
    is_s[sampled] = True
    keys[is_s] = other[: len(sampled)]
    keys[~is_s] = np.concatenate([low, other[len(sampled):]])
    return keys


@pytest.mark.parametrize("q,r,n_crowd", [(16, 5, 10000), (20, 5, 12000)])
@pytest.mark.parametrize("no_dup", [True, False])
def test_sqf_build_crowded_top_bits(q, r, n_crowd, no_dup):
    """10000 / 12000 distinct keys sharing the top 8 fingerprint bits."""
    say("sqf crowded top bits", q, r, n_crowd, no_dup)
    keys = _crowded(61, q, r, n_crowd, 20000)
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=no_dup)
    assert (f.table_bytes() == sqf_canon(keys, q, r, no_dup)).all()


def test_sqf_sorted_lookups_crowded_top_bits():
    q, r = 20, 5
    keys = _crowded(62, q, r, 12000, 100000)
    table = sqf_canon(keys, q, r, True)
    f = G.SQF(q, r)
    f.load_table(table)
    probe = keys  # sorted lookups sort by the top 16 fingerprint bits
    want = O.sqf_lookup_keys(table, q, r, probe)
    assert (sqf_lookup(f, probe, sorted=True) == want).all()


@pytest.mark.parametrize("q,r,n_crowd", [(20, 8, 10000), (24, 10, 6000)])
def test_rsqf_build_crowded_top_bits(q, r, n_crowd):
    """Distinct keys sharing the top 8 fingerprint bits, 28-bit and 34-bit
    fingerprints."""
    keys = _crowded(64, q, r, n_crowd, 50000)
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys))
    assert (f.dump() == _rsqf_canon(keys, q, r)).all()


@pytest.mark.parametrize("n", [1 << 16, 1 << 18])
def test_sqf_build_zipf_small_keys(n):
    """Zipf ranks as keys (small integers, many repeated): their fingerprints
    crowd a few top-bit values."""
    q, r = 20, 5
    keys = bench_zipf(n)
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=True)
    assert (f.table_bytes() == sqf_canon(keys, q, r, True)).all()


def bench_zipf(n, s=1.1, universe=1 << 28, seed=7):
    rng = np.random.default_rng(seed)
    out = np.empty(0, np.uint64)
    while len(out) < n:
        z = rng.zipf(s, 2 * (n - len(out)) + 16).astype(np.uint64)
        out = np.concatenate([out, z[z <= universe]])
    return (out[:n] - 1).astype(np.uint32)


# The partition on the top 8 bits alone serves up to 256 * 7600 keys; larger
# sorts partition again within every segment (2^b2 groups each) before the
# LDS sort of each group.
SEG1_MAX = 256 * 7600


@pytest.fixture
def fresh_stream():
    """A stream of its own: after crowded sorts a stream's next sorts skip
    the partition, so tests of the partition path start on a new one."""
    st = G.Stream()
    yield st
    st.destroy()


@pytest.mark.parametrize("n", [SEG1_MAX, SEG1_MAX + 1, 2600000])
def test_sqf_build_at_the_second_partition(n, fresh_stream):
    """Sizes at the bound of the second partition, 28-bit fingerprints."""
    q, r = 23, 5
    say("sqf build second partition", n)
    keys = O.mt19937(71, n)
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=False, stream=fresh_stream)
    assert (f.table_bytes() == sqf_canon(keys, q, r, False)).all()


@pytest.mark.parametrize("q,r,n", [(24, 10, 2600000), (23, 11, 4200000)])
def test_rsqf_build_at_the_second_partition(q, r, n, fresh_stream):
    """34-bit fingerprints: the groups sort their low 32 bits."""
    say("rsqf build second partition", q, r, n)
    keys = O.mt19937(72, n)
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys), stream=fresh_stream)
    assert (f.dump() == _rsqf_canon(keys, q, r)).all()


@pytest.mark.parametrize("kind", ["hot", "crowded", "zipf"])
def test_sqf_build_skewed_past_the_second_partition(kind):
    """Skewed sorts of more than 256 * 7600 keys: one key 20001 times, 12000
    distinct keys under one top-8-bit value, Zipf ranks as keys."""
    q, r = 24, 5
    n = 2300000
    say("sqf skewed second partition", kind)
    if kind == "hot":
        keys = _with_hot(73, n, 20000)
    elif kind == "crowded":
        keys = _crowded(74, q, r, 12000, n - 12000)
    else:
        keys = bench_zipf(n)
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=True)
    assert (f.table_bytes() == sqf_canon(keys, q, r, True)).all()


def test_rsqf_insert_and_delete_batches_past_the_second_partition(fresh_stream):
    q, r = 24, 10
    say("rsqf insert/delete second partition")
    base = O.mt19937(75, 4000000)
    f = G.RSQF(q, r)
    f.bulk_build(dev(base[:1500000]), stream=fresh_stream)
    f.insertGPU(dev(base[1500000:]), stream=fresh_stream)
    assert (f.dump() == _rsqf_canon(base, q, r)).all()
    gone = base[:2200000]
    f.deleteGPU(dev(gone), stream=fresh_stream)
    assert (f.dump() == _rsqf_canon(base[2200000:], q, r)).all()


def test_crowded_then_uniform_sorts_on_one_stream(fresh_stream):
    """After a crowded sort the stream's next sorts take the chunked passes
    for a while, then probe the partition again: every table stays exact."""
    q, r = 20, 5
    say("crowded then uniform on one stream")
    crowded = bench_zipf(1 << 17)
    f = G.SQF(q, r)
    f.bulk_build(dev(crowded), NoDuplicates=True, stream=fresh_stream)
    assert (f.table_bytes() == sqf_canon(crowded, q, r, True)).all()
    for i in range(7):
        keys = O.mt19937(80 + i, 200000)
        f.bulk_build(dev(keys), NoDuplicates=True, stream=fresh_stream)
        assert (f.table_bytes() == sqf_canon(keys, q, r, True)).all()
