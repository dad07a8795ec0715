"""SQF slot widths the reference GPU code does not support ((r+3) % 8 != 0,
README.md:32-33), through the C ABI: the table is the S slots' (r+3)-bit
elements packed MSB-first into 32-bit words, the CPU_version's order
(quotientFilter.cpp:114-130).

Expected tables are the oracle's end states (kept one element per
ceil((r+3)/8) bytes, pinned to the compiled reference CPU_version at these
widths by tests/test_oracle.py) packed with oracle.sqf_pack, which reproduces
the reference's own words; with the FNV hash the GPU tables are compared with
those words directly."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from conftest import golden
from gpu_helpers import dev, fresh_keys, require_gpu, say, sqf_lookup

pytestmark = pytest.mark.gpu

PACKED_R = [1, 2, 4, 6, 7, 8, 9, 10, 11, 12]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def words(qf):
    return qf.table_bytes().view(np.uint32)


def packed(table, q, r):
    return O.sqf_pack(O.sqf_decode(table, q, r), q, r)


def canon(keys, q, r, no_dup):
    st, t = O.sqf_canon(O.fingerprints(keys, q, r), q, r, no_dup)
    assert st == O.OK
    return t


@pytest.mark.parametrize("name", ["sqf_fnv_q12_r6_a0.7.npz", "sqf_fnv_q12_r7_a0.75.npz", "sqf_fnv_q12_r8_a0.8.npz",
                                  "sqf_fnv_q11_r9_a0.8.npz", "sqf_fnv_q11_r10_a0.85.npz",
                                  "sqf_fnv_q10_r11_a0.85.npz", "sqf_fnv_q10_r12_a0.9.npz"])
def test_fnv_tables_are_the_reference_cpu_words(name):
    """FNV-hashed build == the reference CPU_version's insertItem end state,
    word for word; lookups == its mayContain; the CPU-layout image is its
    table."""
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    say("sqf packed vs reference words", name)
    G.set_sqf_hash(G.HASH_FNV)
    try:
        qf = G.SQF(q, r)
        assert qf.h.bytesPerElement == 0
        assert G.lib().gqf_sqf_table_bytes(q, r) == 4 * O.sqf_packed_words(q, r)
        qf.bulk_build(dev(g["keys"]), NoDuplicates=True)
        w = words(qf)
        ref = g["words"].astype(np.uint32)
        assert (w[: len(ref)] == ref).all() and not w[len(ref):].any()
        assert (sqf_lookup(qf, g["lookup_keys"]) == g["lookup_ret"]).all()
        img = G.image_save(qf, G.IMAGE_SQF_CPU)
        assert (np.frombuffer(img[64:], dtype="<u4") == ref).all()
        h = G.SQF(q, r)
        G.image_load(h, img)
        assert (words(h) == w).all()
        # the same keys inserted in batches give the same table
        f2 = G.SQF(q, r)
        k = g["keys"]
        for i in range(0, len(k), 700):
            f2.insert(dev(k[i:i + 700]))
        assert (words(f2) == w).all()
    finally:
        G.set_sqf_hash(G.HASH_AP)


@pytest.mark.parametrize("r", PACKED_R)
@pytest.mark.parametrize("no_dup", [True, False])
def test_build_and_lookup_equal_oracle(r, no_dup):
    q = min(16, 32 - r)
    say("sqf packed build", q, r, no_dup)
    n = O.key_count(0.85, q)
    keys = O.mt19937(3, n)
    if not no_dup:
        keys[n // 2:] = keys[: n - n // 2]
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys), NoDuplicates=no_dup)
    t = canon(keys, q, r, no_dup)
    assert (words(qf) == packed(t, q, r)).all()
    probe = np.concatenate([keys, fresh_keys(4, n, keys)])
    want = O.sqf_lookup_keys(t, q, r, probe)
    assert (sqf_lookup(qf, probe) == want).all()
    assert (sqf_lookup(qf, probe, sorted=True) == want).all()


@pytest.mark.parametrize("q,r", [(14, 8), (13, 6), (12, 11), (15, 3)])
def test_insert_delete_merge_equal_oracle(q, r):
    say("sqf packed insert/delete/merge", q, r)
    n = O.key_count(0.85, q)
    keys = O.mt19937(5, n)
    half = O.key_count(0.5, q)
    qf = G.SQF(q, r)
    qf.bulk_build(dev(keys[:half]), NoDuplicates=True)
    t = canon(keys[:half], q, r, True)
    step = 1 << (q - 3)
    for i in range(half, n, step):
        batch = keys[i:i + step]
        qf.insert(dev(batch))
        assert O.sqf_insert_keys(t, q, r, batch) == O.OK
    assert (words(qf) == packed(t, q, r)).all()
    d = np.concatenate([keys[: n // 5], O.mt19937(6, 300)])
    qf.superclusterDeletes(dev(d))
    O.sqf_delete_keys(t, q, r, d)
    assert (words(qf) == packed(t, q, r)).all()
    # insertViaMerge keeps duplicates (NoDuplicates=False): the stored
    # multiset (read back from the oracle's table) plus the new fingerprints
    extra = keys[n // 5: n // 5 + 200]
    qf.insertViaMerge(dev(extra), NoDuplicates=False)
    want = np.concatenate([O.sqf_extract(t, q, r), O.fingerprints(extra, q, r)])
    st, t2 = O.sqf_canon(want, q, r, False)
    assert st == O.OK
    assert (words(qf) == packed(t2, q, r)).all()
    # mergeTwoFilters of two packed filters
    a, b = G.SQF(q, r), G.SQF(q, r)
    ka, kb = keys[: n // 3], keys[n // 4: n // 2]
    a.bulk_build(dev(ka), NoDuplicates=True)
    b.bulk_build(dev(kb), NoDuplicates=True)
    a.mergeTwoFilters(b, NoDuplicates=True)
    assert (words(a) == packed(canon(np.concatenate([ka, kb]), q, r, True), q, r)).all()
    assert (words(b) == packed(canon(kb, q, r, True), q, r)).all()


def test_capacity_error_leaves_a_packed_table_unchanged():
    q, r = 10, 8
    qf = G.SQF(q, r)
    keys = O.mt19937(8, 800)
    qf.bulk_build(dev(keys), NoDuplicates=True)
    before = words(qf).copy()
    with pytest.raises(G.CapacityError):
        qf.bulk_build(dev(O.mt19937(9, 4000)), NoDuplicates=False)
    assert (words(qf) == before).all()
    with pytest.raises(G.CapacityError):
        qf.insert(dev(O.mt19937(10, 4000)))
    assert (words(qf) == before).all()


def test_full_size_round_trip_and_false_positives():
    """q=24, r=8 (11-bit elements, 23.1 MB): every inserted key is found, a
    fresh key is reported present exactly when its fingerprint is stored (the
    quotient filter's only false positives), and a delete of every inserted
    key empties the table.  (At 32-bit fingerprints the reference AP hash is
    not uniform: about 3x the m/2^32 false positives of a uniform hash, so
    the count is checked exactly rather than against a binomial band.)"""
    q, r = 24, 8
    n = O.key_count(0.8, q)
    keys = G.mt19937(1, n)
    say("sqf packed full size", q, r)
    qf = G.SQF(q, r)
    d = dev(keys)
    qf.bulk_build(d, NoDuplicates=True)
    out = G.DeviceArray.empty(n)
    qf.lookup(d, out)
    assert (out.to_host() != G.NOT_FOUND).all()
    fresh = G.mt19937(2, 1 << 22)
    fresh = fresh[~np.isin(fresh, keys)]
    hits = sqf_lookup(qf, fresh) != G.NOT_FOUND
    stored = np.unique(O.fingerprints(keys, q, r))
    assert (hits == np.isin(O.fingerprints(fresh, q, r), stored)).all()
    qf.superclusterDeletes(d)
    assert not words(qf).any()


def test_batches_refuse_packed_widths():
    f = G.SQF(12, 8)
    keys = dev(O.mt19937(1, 1000))
    with pytest.raises(G.GQFError) as e:
        G.SQFBatch([f], keys, [0, 1000], no_dup=True)
    assert e.value.status == G.ERR_UNSUPPORTED
