"""SQF on the GPU through the C ABI, bit-exact against the oracle.

Tables must equal the canonical layout byte for byte; lookups must equal the
oracle's restatement of hashAndLookUp (quotientFilter.cu:320-357)."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from gpu_helpers import dev, fresh_keys, require_gpu, say, sqf_canon, sqf_lookup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


BUILD_CASES = [(10, 5, 0.6), (14, 5, 0.8), (16, 5, 0.95), (12, 13, 0.8), (16, 13, 0.9), (20, 5, 0.8)]


@pytest.mark.parametrize("q,r,alpha", BUILD_CASES)
@pytest.mark.parametrize("no_dup", [True, False])
@pytest.mark.parametrize("method", [G.BUILD_SEGMENTED_LAYOUTS, G.BUILD_PARALLEL_MERGING, G.BUILD_SEQUENTIAL_SHIFTS])
def test_build_equals_canonical(q, r, alpha, no_dup, method):
    say("sqf build", q, r, alpha, no_dup, method)
    n = O.key_count(alpha, q)
    keys = O.mt19937(1, n)
    if not no_dup:  # make real duplicates appear
        keys[n // 2:] = keys[: n - n // 2]
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys), NoDuplicates=no_dup, method=method)
    want = sqf_canon(keys, q, r, no_dup)
    got = qf.table_bytes()
    assert (got == want).all(), f"first diff at byte {np.flatnonzero(got != want)[:5]}"


@pytest.mark.parametrize("q,r,alpha", [(12, 5, 0.6), (16, 5, 0.9), (14, 13, 0.9), (20, 5, 0.8)])
@pytest.mark.parametrize("sorted_", [False, True])
def test_lookup_matches_oracle(q, r, alpha, sorted_):
    say("sqf lookup", q, r, alpha, sorted_)
    n = O.key_count(alpha, q)
    keys = O.mt19937(1, n)
    table = sqf_canon(keys, q, r, True)
    qf = G.SQF(q, r)
    qf.load_table(table)
    probe = np.concatenate([keys, fresh_keys(2, n, keys)])
    want = O.sqf_lookup_keys(table, q, r, probe)
    got = sqf_lookup(qf, probe, sorted=sorted_)
    assert (got == want).all()
    assert (got[:n] != G.NOT_FOUND).all()  # no false negatives


def test_lookup_on_reference_insert_table():
    # a table produced by the faithful restatement of the reference insert loop
    q, r = 13, 5
    keys = O.mt19937(5, O.key_count(0.85, q))
    t = np.zeros(O.sqf_table_bytes(q, r), np.uint8)
    assert O.sqf_insert_keys(t, q, r, keys) == O.OK
    qf = G.SQF(q, r)
    qf.load_table(t)
    probe = np.concatenate([keys, O.mt19937(6, 4000)])
    assert (sqf_lookup(qf, probe) == O.sqf_lookup_keys(t, q, r, probe)).all()


@pytest.mark.parametrize("q,r", [(14, 5), (16, 13), (18, 5)])
def test_insert_batches_equal_sequential_reference_insert(q, r):
    say("sqf insert", q, r)
    n = O.key_count(0.85, q)
    keys = O.mt19937(3, n)
    half = O.key_count(0.5, q)
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys[:half]), NoDuplicates=True)
    t = sqf_canon(keys[:half], q, r, True)
    step = 1 << (q - 4)
    for i in range(half, n, step):
        batch = keys[i:i + step]
        qf.insert(dev(batch))
        assert O.sqf_insert_keys(t, q, r, batch) == O.OK
    assert (qf.table_bytes() == t).all()


def test_insert_dedups_against_dup_build():
    q, r = 12, 5
    keys = O.mt19937(9, 2000)
    base = np.concatenate([keys[:800], keys[:300]])  # duplicates stored (Dup build)
    qf = G.SQF(q, r)
    qf.bulk_build(dev(base), NoDuplicates=False)
    new = np.concatenate([keys[200:1200], keys[1000:1100]])
    qf.insert(dev(new))
    # per distinct fingerprint: max(stored count, [present in batch])
    fb = O.fingerprints(base, q, r)
    fn = O.fingerprints(new, q, r)
    vals, cnt = np.unique(fb, return_counts=True)
    extra = np.setdiff1d(np.unique(fn), vals)
    want_multiset = np.concatenate([np.repeat(vals, cnt), extra])
    st, want = O.sqf_canon(want_multiset, q, r, False)
    assert (qf.table_bytes() == want).all()


@pytest.mark.parametrize("q,r,frac", [(12, 5, 0.1), (16, 5, 0.5), (14, 13, 0.3), (18, 5, 0.1)])
def test_delete_equals_reference_restatement(q, r, frac):
    say("sqf delete", q, r, frac)
    n = O.key_count(0.8, q)
    keys = O.mt19937(1, n)
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys), NoDuplicates=True)
    t = sqf_canon(keys, q, r, True)
    d = np.concatenate([keys[: int(n * frac)], O.mt19937(77, 500)])
    qf.superclusterDeletes(dev(d))
    O.sqf_delete_keys(t, q, r, d)
    assert (qf.table_bytes() == t).all()
    got = sqf_lookup(qf, keys)
    assert (got == O.sqf_lookup_keys(t, q, r, keys)).all()


def test_delete_multiset_on_dup_table():
    q, r = 12, 5
    keys = O.mt19937(4, 1500)
    base = np.concatenate([keys, keys[:400], keys[:100]])
    qf = G.SQF(q, r)
    qf.bulk_build(dev(base), NoDuplicates=False)
    t = sqf_canon(base, q, r, False)
    d = np.concatenate([keys[:50], keys[:50], keys[:250], keys[1400:]])
    qf.delete(dev(d))
    O.sqf_delete_keys(t, q, r, d)
    assert (qf.table_bytes() == t).all()


def test_skewed_tile_insert_and_delete():
    # every fingerprint in the first 2048 quotients: one cluster several times
    # longer than its quotient range, spanning many rebuild sub-tiles (long
    # fix-up lists, owned ranges far from their quotients)
    q, r = 14, 5
    pool = O.mt19937(21, 120000)
    fq = O.fingerprints(pool, q, r) >> r
    skew = pool[fq < 2048][:6000]
    assert len(skew) == 6000
    say("sqf skewed tile", q, r, len(skew))
    qf = G.SQF(q, r)
    qf.bulk_build(dev(skew[:3000]), NoDuplicates=True)
    qf.insert(dev(skew[3000:]))
    want = sqf_canon(skew, q, r, True)
    assert (qf.table_bytes() == want).all()
    d = skew[1000:4000]
    qf.superclusterDeletes(dev(d))
    O.sqf_delete_keys(want, q, r, d)
    assert (qf.table_bytes() == want).all()


@pytest.mark.parametrize("q", [12, 16])
def test_delete_heavy_duplicates(q):
    # many copies per fingerprint (runs of equal entries inside one thread's
    # slots and across threads); deletes remove up to, and past, every copy
    r = 5
    keys = O.mt19937(31, O.key_count(0.05, q))
    reps = np.array([1, 2, 3, 5, 8])[np.arange(len(keys)) % 5]
    base = np.repeat(keys, reps)
    say("sqf delete duplicates", q, len(base))
    qf = G.SQF(q, r)
    qf.bulk_build(dev(base), NoDuplicates=False)
    t = sqf_canon(base, q, r, False)
    assert (qf.table_bytes() == t).all()
    d = np.concatenate([np.repeat(keys[::2], 2), np.repeat(keys[1::3], 7), keys[::5]])
    qf.delete(dev(d))
    O.sqf_delete_keys(t, q, r, d)
    assert (qf.table_bytes() == t).all()


@pytest.mark.parametrize("no_dup", [True, False])
def test_insert_via_merge(no_dup):
    q, r = 15, 5
    keys = O.mt19937(12, O.key_count(0.8, q))
    old, new = keys[:15000], np.concatenate([keys[15000:], keys[:2000]])
    qf = G.SQF(q, r)
    qf.bulk_build(dev(old), NoDuplicates=no_dup)
    qf.insertViaMerge(dev(new), NoDuplicates=no_dup)
    want = sqf_canon(np.concatenate([old, new]), q, r, no_dup)
    assert (qf.table_bytes() == want).all()


@pytest.mark.parametrize("no_dup", [True, False])
def test_merge_two_filters(no_dup):
    q, r = 14, 5
    keys = O.mt19937(13, O.key_count(0.8, q))
    a, b = keys[:6000], keys[5000:]
    f1, f2 = G.SQF(q, r), G.SQF(q, r)
    f1.bulk_build(dev(a), NoDuplicates=True)
    f2.bulk_build(dev(b), NoDuplicates=True)
    before2 = f2.table_bytes()
    f1.mergeTwoFilters(f2, NoDuplicates=no_dup)
    fa = np.unique(O.fingerprints(a, q, r))
    fb = np.unique(O.fingerprints(b, q, r))
    st, want = O.sqf_canon(np.concatenate([fa, fb]), q, r, no_dup)
    assert (f1.table_bytes() == want).all()
    assert (f2.table_bytes() == before2).all()


def test_edge_cases():
    q, r = 10, 5
    qf = G.SQF(q, r)
    # empty input builds an empty table
    qf.bulk_build(dev(np.zeros(0, np.uint32)), n=0)
    assert (qf.table_bytes() == 0).all()
    # 0xFFFFFFFF is an ordinary key (the reference's queue treats it as a sentinel)
    k = np.array([0xFFFFFFFF, 0, 1, 0xFFFFFFFE], np.uint32)
    qf.insert(dev(k))
    assert (qf.table_bytes() == sqf_canon(k, q, r, True)).all()
    assert (sqf_lookup(qf, k) != G.NOT_FOUND).all()
    qf.delete(dev(k[:1]))
    assert (qf.table_bytes() == sqf_canon(k[1:], q, r, True)).all()
    # deleting absent keys changes nothing
    before = qf.table_bytes()
    qf.delete(dev(O.mt19937(99, 100)))
    assert (qf.table_bytes() == before).all()


def test_capacity_error_handling():
    q, r = 8, 5
    S = O.sqf_num_slots(q)
    keys = O.mt19937(1, 200)
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys))
    before = qf.table_bytes()
    too_many = O.mt19937(7, S + 60)  # more distinct fingerprints than slots
    assert len(np.unique(O.fingerprints(too_many, q, r))) > S
    with pytest.raises(G.CapacityError):
        qf.insert(dev(too_many))
    assert (qf.table_bytes() == before).all()
    qf.delete(dev(keys[:50]))  # still usable after the error
    before = qf.table_bytes()
    assert (before == sqf_canon(keys[50:], q, r, True)).all()
    # a build that does not fit is rejected before any write
    with pytest.raises(G.CapacityError):
        qf.bulk_build(dev(too_many), NoDuplicates=False)
    assert (qf.table_bytes() == before).all()


def test_deterministic_rebuild():
    q, r = 18, 5
    keys = O.mt19937(31, O.key_count(0.9, q))
    a, b = G.SQF(q, r), G.SQF(q, r)
    a.bulk_build(dev(keys))
    b.bulk_build(dev(keys[::-1].copy()))
    assert (a.table_bytes() == b.table_bytes()).all()


def test_large_rebuilds_equal_canonical():
    """q=24 (up to 13 M entries, more than 1024 rebuild chunks, so the chunk
    scan combines several workgroups): build, two merge inserts of 2^21
    keys and a delete, each against the oracle's canonical layout."""
    q, r = 24, 5
    n = O.key_count(0.8, q)
    keys = O.mt19937(41, n)
    half = O.key_count(0.5, q)
    say("sqf large rebuilds", q, r, n)
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys[:half]), NoDuplicates=True)
    assert (qf.table_bytes() == sqf_canon(keys[:half], q, r, True)).all()
    end = half
    for _ in range(2):
        qf.insert(dev(keys[end:end + (1 << 21)]))
        end += 1 << 21
    assert (qf.table_bytes() == sqf_canon(keys[:end], q, r, True)).all()
    # NoDup table: one entry per distinct fingerprint; deleting a key removes
    # its fingerprint's entry
    d = keys[: 1 << 20]
    qf.delete(dev(d))
    left = np.setdiff1d(np.unique(O.fingerprints(keys[:end], q, r)), O.fingerprints(d, q, r))
    st, want = O.sqf_canon(left, q, r, True)
    assert st == O.OK and (qf.table_bytes() == want).all()


def test_full_size_q26_properties():
    """Full size (q=26, r=5, fill 0.8): checks that need no oracle
    at this size -- no false negatives, FPR band, sorted == unsorted, extract
    round trip count, and a delete+lookup round trip."""
    q, r = 26, 5
    n = O.key_count(0.8, q)
    say("sqf full size", q, r, n)
    keys = O.mt19937(1, n)
    dk = dev(keys)
    qf = G.SQF(q, r)
    qf.bulkBuildSegmentedLayouts(dk, NoDuplicates=True)
    m = 1 << 20
    probe = np.concatenate([keys[: m // 2], fresh_keys(2, m // 2, keys)])
    dp = dev(probe)
    o1, o2 = G.DeviceArray.empty(m), G.DeviceArray.empty(m)
    qf.launchUnsortedLookups(dp, o1)
    qf.launchSortedLookups(dp, o2)
    u, s = o1.to_host(), o2.to_host()
    assert (u == s).all()
    assert (u[: m // 2] != G.NOT_FOUND).all()
    # exact on fingerprints: hit <=> the probe's fingerprint is stored (the FPR
    # is then the reference AP hash's own collision rate at 31 bits)
    fk = O.fingerprints(keys, q, r)
    assert ((u != G.NOT_FOUND) == np.isin(O.fingerprints(probe, q, r), fk)).all()
    nu = len(np.unique(fk))
    # occupancy: number of non-empty slots == distinct fingerprints
    t = qf.table_bytes()
    assert int(((t & 7) != 0).sum()) == nu
    # delete the first 2^20 keys, then they are absent unless a survivor shares the fingerprint
    qf.delete(dev(keys[:m]))
    t2 = qf.table_bytes()
    fk = O.fingerprints(keys, q, r)
    remaining = np.setdiff1d(np.unique(fk), np.unique(fk[:m]))
    assert int(((t2 & 7) != 0).sum()) == len(remaining)


def test_concurrent_streams_from_threads():
    # four host threads, each with its own stream and filters, calling at once:
    # every result must equal the single-threaded oracle result
    import threading

    say("sqf/rsqf concurrent streams", 4)
    jobs = []
    for t in range(4):
        q = 14 + t
        keys = O.mt19937(40 + t, O.key_count(0.7, q))
        jobs.append((q, keys))
    errors = []

    def work(q, keys):
        try:
            s = G.Stream()
            qf = G.SQF(q, 5)
            for rep in range(3):
                qf.bulk_build(dev(keys), NoDuplicates=True, stream=s)
                qf.insert(dev(O.mt19937(99, 500)), stream=s)
                qf.superclusterDeletes(dev(keys[:300]), stream=s)
            f = G.RSQF(q, 8)
            f.bulk_build(dev(keys), stream=s)
            f.insertGPU(dev(O.mt19937(98, 400)), stream=s)
            table, dumped = qf.table_bytes(), f.dump()
            s.destroy()
            jobs_out[q] = (table, dumped)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    jobs_out = {}
    th = [threading.Thread(target=work, args=j) for j in jobs]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    for q, keys in jobs:
        want = sqf_canon(np.concatenate([keys, O.mt19937(99, 500)]), q, 5, True)
        O.sqf_delete_keys(want, q, 5, keys[:300])
        got_t, got_r = jobs_out[q]
        assert (got_t == want).all()
        st, rw = O.rsqf_canon(O.fingerprints(np.concatenate([keys, O.mt19937(98, 400)]), q, 8), q, 8)
        assert (got_r == rw).all()


def _insert_both_paths(q, r, base, batches, no_dup=True, paths=(G.INSERT_LOCAL, G.INSERT_REBUILD)):
    """Apply the insert batches to filters built from base, one per insert
    path (localized, full rebuild); the tables (and statuses) must be equal
    after every batch."""
    out = []
    for path in paths:
        G.set_sqf_insert_path(path)
        try:
            f = G.SQF(q, r)
            if len(base):
                f.bulk_build(dev(base), NoDuplicates=no_dup)
            tables, status = [], []
            for b in batches:
                try:
                    f.insert(dev(b))
                    status.append(G.OK)
                except G.CapacityError:
                    status.append(G.ERR_CAPACITY)
                tables.append(f.table_bytes())
            out.append((tables, status))
            f.destroy()
        finally:
            G.set_sqf_insert_path(G.INSERT_AUTO)
    tl, sl = out[0]
    for p, (tr, sr) in zip(paths[1:], out[1:]):
        assert sl == sr, ("path", p)
        for i, (a, b) in enumerate(zip(tl, tr)):
            assert (a == b).all(), ("path", p, "batch", i, np.flatnonzero(a != b)[:5])
    return tl, sl


@pytest.mark.parametrize("q,r,a0,a1,nb", [(14, 5, 0.5, 0.85, 6), (16, 13, 0.6, 0.9, 5), (11, 21, 0.3, 0.8, 4),
                                          (18, 5, 0.05, 0.5, 3), (12, 5, 0.0, 0.6, 4)])
def test_insert_local_equals_rebuild_and_reference(q, r, a0, a1, nb):
    """Localized inserts (regions around each key) equal the full rebuild and
    the reference's sequential insert, batch by batch."""
    say("sqf localized insert", q, r, a0, a1)
    n0, n1 = O.key_count(a0, q), O.key_count(a1, q)
    keys = O.mt19937(51 + q, n1)
    step = max(1, (n1 - n0) // nb)
    batches = [keys[i:i + step] for i in range(n0, n1, step)]
    tables, status = _insert_both_paths(q, r, keys[:n0], batches)
    assert all(s == G.OK for s in status)
    t = sqf_canon(keys[:n0], q, r, True) if n0 else np.zeros(O.sqf_table_bytes(q, r), np.uint8)
    for b, got in zip(batches, tables):
        assert O.sqf_insert_keys(t, q, r, b) == O.OK
        assert (got == t).all()


def test_insert_local_duplicates_hot_keys_and_edges():
    """Duplicates in a batch and against a Dup-mode table, keys whose quotients
    sit at both ends of the table, and 3000 copies of one key (one region
    over thousands of slots, of which one entry is new)."""
    q, r = 13, 5
    S = G.SQF.num_slots(q)
    say("sqf localized insert edges", q, r)
    base = O.mt19937(61, O.key_count(0.4, q))
    base = np.concatenate([base, base[:300]])  # stored twice (Dup build)
    fp = O.fingerprints(np.arange(200000, dtype=np.uint32), q, r)
    qs = fp >> r
    low = np.arange(200000, dtype=np.uint32)[qs < 8][:20]
    high = np.arange(200000, dtype=np.uint32)[qs >= (1 << q) - 8][:20]
    hot = np.full(3000, 12345, np.uint32)
    batches = [np.concatenate([base[:100], base[:100], O.mt19937(62, 200), O.mt19937(62, 200)]),
               np.concatenate([low, high]), hot, O.mt19937(63, 500)]
    _insert_both_paths(q, r, base, batches, no_dup=False)


def test_insert_local_near_capacity():
    """A batch that does not fit: the localized path hands over to the
    rebuild, which reports the capacity error and leaves the table as it was."""
    q, r = 12, 5
    S = G.SQF.num_slots(q)
    base = O.mt19937(71, int(0.97 * S))
    extra = O.mt19937(72, 400)
    say("sqf localized insert capacity", q, r)
    tables, status = _insert_both_paths(q, r, base, [extra[:50], extra])
    assert status[-1] == G.ERR_CAPACITY and (tables[-1] == tables[-2]).all()


def test_insert_local_long_region_falls_back():
    """70,000 copies of one key: the localized path's region would pass 2^16
    slots, so the call runs the rebuild; the result is the same."""
    q, r = 18, 5
    say("sqf localized insert fallback", q, r)
    base = O.mt19937(81, O.key_count(0.3, q))
    _insert_both_paths(q, r, base, [np.full(70000, 777, np.uint32), O.mt19937(82, 3000)])


@pytest.mark.parametrize("q,r,a0,a1,nb", [(20, 5, 0.0005, 0.004, 3), (3, 29, 0.0, 0.8, 3), (9, 13, 0.2, 0.95, 3),
                                          (11, 21, 0.4, 0.9, 2), (21, 5, 0.7, 0.9, 2)])
def test_insert_sparse_tiny_and_wide_tables(q, r, a0, a1, nb):
    """Sparse tables (most 4096-slot decode regions empty), a table smaller
    than one region, every element width, and full ones: both insert paths
    against the reference's sequential insert."""
    say("sqf insert sparse/tiny/wide", q, r, a0, a1)
    n0, n1 = O.key_count(a0, q), O.key_count(a1, q)
    keys = O.mt19937(91 + q, n1)
    step = max(1, (n1 - n0) // nb)
    batches = [keys[i:i + step] for i in range(n0, n1, step)]
    tables, status = _insert_both_paths(q, r, keys[:n0], batches)
    assert all(s == G.OK for s in status)
    t = sqf_canon(keys[:n0], q, r, True) if n0 else np.zeros(O.sqf_table_bytes(q, r), np.uint8)
    for b, got in zip(batches, tables):
        assert O.sqf_insert_keys(t, q, r, b) == O.OK
        assert (got == t).all()


def test_insert_long_pending_chains():
    """12,000 fingerprints in the first 2048 quotients: at the second decode
    region's start more than 1024 occupied quotients still wait for their
    runs, so decoding walks clusters for some quotients; hot keys put more
    than a sub-tile of new entries between two stored ones."""
    q, r = 14, 5
    pool = O.mt19937(23, 200000)
    fq = O.fingerprints(pool, q, r) >> r
    skew = np.unique(pool[fq < 2048])[:12000]
    assert len(skew) == 12000
    say("sqf insert pending chains", q, r)
    hot = np.concatenate([np.full(5000, skew[7], np.uint32), skew[6000:9000], np.full(3000, 4242, np.uint32)])
    batches = [skew[4000:6000], hot, skew[9000:]]
    tables, status = _insert_both_paths(q, r, skew[:4000], batches, no_dup=False)
    assert all(s == G.OK for s in status)
    t = sqf_canon(skew[:4000], q, r, False)
    for b, got in zip(batches, tables):
        assert O.sqf_insert_keys(t, q, r, b) == O.OK
        assert (got == t).all()


@pytest.mark.parametrize("no_dup", [True, False])
def test_insert_via_merge_heavy_duplicates(no_dup):
    """insertViaMerge on a table holding every key three times, with new keys
    repeated four times: keep-all keeps every copy, no-dup one of each."""
    q, r = 13, 5
    keys = O.mt19937(94, 1500)
    base = np.repeat(keys, 3)
    new = np.concatenate([np.repeat(keys[::7], 4), O.mt19937(95, 900)])
    f = G.SQF(q, r)
    f.bulk_build(dev(base), NoDuplicates=False)
    f.insertViaMerge(dev(new), NoDuplicates=no_dup)
    want = sqf_canon(np.concatenate([base, new]), q, r, no_dup)
    assert (f.table_bytes() == want).all()
