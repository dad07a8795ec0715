"""The oracle pinned against the reference's own outputs (golden fixtures
written by tests/golden/make_golden.py from the compiled reference CPU code),
against closed forms, and against brute force -- before anything is checked
against it."""
import numpy as np
import pytest

from conftest import golden


def test_mt19937_matches_reference_stream(orc):
    g = golden("mt19937.npz")
    assert (orc.mt19937(1, 4096) == g["seed1"]).all()
    assert (orc.mt19937(7, 1024) == g["seed7"]).all()
    # C++ standard known answer: 10000th output of mt19937 default seed 5489
    assert int(orc.mt19937(5489, 1, skip=9999)[0]) == 4123659995


def test_key_count_rule():
    import oracle

    # (int)(float alpha * 2^q) as the reference drivers compute it
    assert oracle.key_count(0.8, 26) == 53687092
    assert oracle.key_count(0.9, 26) == 60397976
    assert oracle.key_count(0.6, 20) == 629145
    assert oracle.key_count(0.7, 26) == 46976204


def test_ap_and_fnv_hash_golden(orc):
    g = golden("hash.npz")
    for i, bits in enumerate(g["bits"]):
        bits = int(bits)
        ap = orc.fingerprints(g["keys"], bits - 5, 5, 0)
        fnv = orc.fingerprints(g["keys"], bits - 5, 5, 1)
        assert (ap == g["ap"][i]).all(), bits
        assert (fnv == g["fnv"][i]).all(), bits


def _fmix64_inverse(h):
    """Inverse of the MurmurHash3 finalizer, from its definition (xor-shift by
    33 is its own inverse on 64 bits; the multipliers are odd)."""
    M = (1 << 64) - 1
    inv1 = pow(0xFF51AFD7ED558CCD, -1, 1 << 64)
    inv2 = pow(0xC4CEB9FE1A85EC53, -1, 1 << 64)
    h ^= h >> 33
    h = (h * inv2) & M
    h ^= h >> 33
    h = (h * inv1) & M
    h ^= h >> 33
    return h


def test_wide_fingerprint_rule(orc):
    keys = orc.mt19937(5, 4096)
    lo = np.array([orc.lib().orc_ap_rounds(int(k), 0xAAAAAAAA) for k in keys], np.uint64)
    # q + r == 32: the plain 32-bit AP value
    assert (orc.fingerprints(keys, 27, 5) == lo).all()
    # below 32 bits: the reference modulus
    assert (orc.fingerprints(keys, 20, 5) == lo % np.uint64(1 << 25)).all()
    # q + r > 32: fmix64(AP << 32 | key); inverting the finalizer recovers the input word
    for k, a in zip(keys[:256], lo[:256]):
        h = orc.lib().orc_fmix64((int(a) << 32) | int(k))
        assert _fmix64_inverse(h) == (int(a) << 32) | int(k)
        assert orc.fp_ap(int(k), 26, 8) == h & ((1 << 34) - 1)
    # the published finalizer maps 0 to 0 and is not the identity elsewhere
    assert orc.lib().orc_fmix64(0) == 0 and orc.lib().orc_fmix64(1) != 1


def test_wide_fingerprint_is_uniform(orc):
    # 34-bit fingerprints of 2^22 keys: distinct count matches the birthday
    # bound of a uniform hash, and the top bits are balanced
    n = 1 << 22
    keys = np.unique(orc.mt19937(1, n))  # ~2^11 of the keys repeat; count distinct keys
    k = len(keys)
    fp = orc.fingerprints(keys, 26, 8)
    d = len(np.unique(fp))
    expect = k - k * (k - 1) / 2 / 2.0 ** 34
    assert abs(d - expect) < 6 * np.sqrt(k * k / 2 / 2.0 ** 34)
    top = np.bincount((fp >> np.uint64(26)).astype(np.int64), minlength=256)
    chi2 = ((top - n / 256) ** 2 / (n / 256)).sum()
    assert chi2 < 400  # 255 degrees of freedom


def test_rank_select_golden_and_definition(orc):
    g = golden("rank_select.npz")
    L = orc.lib()
    for j, x in enumerate(g["words"][:512]):
        x = int(x)
        for k in g["ks"]:
            assert L.orc_select(x, int(k)) == g["select"][j, k]
        for i in range(64):
            assert L.orc_rank(x, i) == g["rank"][j, i]
    # definition by brute force on sparse words
    for x in [0, 1, 1 << 63, 0x8000000000000001, 0xF0F0F0F0F0F0F0F0]:
        pos = [i for i in range(64) if (x >> i) & 1]
        for k in range(1, 66):
            want = pos[k - 1] if k <= len(pos) else 0xFFFFFFFF
            assert L.orc_select(x, k) == want


def _geom(orc, q, r):
    return orc.sqf_num_slots(q), orc.sqf_table_bytes(q, r)


def test_sqf_geometry(orc):
    assert orc.sqf_num_slots(26) == 73819750
    assert orc.sqf_table_bytes(26, 5) == 73819750
    assert orc.sqf_table_bytes(18, 13) == 576716
    assert orc.sqf_num_slots(20) == 1153433
    assert orc.lib().orc_sqf_cpu_words(20, 5) == 288358
    assert orc.rsqf_num_blocks(26) == 1048577
    assert orc.rsqf_table_bytes(26, 8) == 84934737


SQF_PACKED_FIXTURES = ["sqf_fnv_q12_r6_a0.7.npz", "sqf_fnv_q12_r7_a0.75.npz", "sqf_fnv_q12_r8_a0.8.npz",
                       "sqf_fnv_q11_r9_a0.8.npz", "sqf_fnv_q11_r10_a0.85.npz", "sqf_fnv_q10_r11_a0.85.npz",
                       "sqf_fnv_q10_r12_a0.9.npz"]


@pytest.mark.parametrize("name", ["sqf_fnv_q12_r5_a0.6.npz", "sqf_fnv_q14_r5_a0.9.npz", "sqf_fnv_q12_r13_a0.8.npz"]
                         + SQF_PACKED_FIXTURES)
def test_sqf_oracle_matches_reference_cpu(orc, name):
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    ref_elems = orc.sqf_decode_words(g["words"], q, r)
    # (1) canonical closed form == reference insertItem end state
    fps = orc.fingerprints(g["keys"], q, r, 1)
    st, can = orc.sqf_canon(fps, q, r, True)
    assert st == orc.OK
    assert (orc.sqf_decode(can, q, r) == ref_elems).all()
    # (2) faithful sequential insert restatement == reference end state
    t = np.zeros(orc.sqf_table_bytes(q, r), np.uint8)
    assert orc.sqf_insert_keys(t, q, r, g["keys"], 1) == orc.OK
    assert (orc.sqf_decode(t, q, r) == ref_elems).all()
    # (3) lookups == reference mayContain
    out = orc.sqf_lookup_keys(can, q, r, g["lookup_keys"], 1)
    assert (out == g["lookup_ret"]).all()


@pytest.mark.parametrize("name", SQF_PACKED_FIXTURES)
def test_sqf_packer_reproduces_reference_words(orc, name):
    """The packed layout of widths (r+3) % 8 != 0 is the CPU_version's word
    layout over the GPU's S slots: packing the oracle's canonical elements
    gives the reference's own words (then zeros: the GPU table is a few words
    longer), and unpacking any words gives back their elements."""
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    ref = g["words"].astype(np.uint32)
    st, can = orc.sqf_canon(orc.fingerprints(g["keys"], q, r, 1), q, r, True)
    assert st == orc.OK
    packed = orc.sqf_pack(orc.sqf_decode(can, q, r), q, r)
    assert len(packed) == orc.sqf_packed_words(q, r) >= len(ref)
    assert (packed[: len(ref)] == ref).all() and not packed[len(ref):].any()
    rnd = orc.mt19937(7, orc.sqf_num_slots(q)) & np.uint32((1 << (r + 3)) - 1)
    assert (orc.sqf_decode_words(orc.sqf_pack(rnd, q, r), q, r) == rnd).all()


def test_sqf_delete_restatement_is_canonical(orc):
    q, r = 13, 5
    keys = orc.mt19937(3, orc.key_count(0.85, q))
    fps = orc.fingerprints(keys, q, r)
    st, can = orc.sqf_canon(fps, q, r, False)  # Dup mode keeps repeated fingerprints
    d = np.concatenate([keys[:1500], keys[:200], orc.mt19937(4, 300)])
    t = can.copy()
    removed = orc.sqf_delete_keys(t, q, r, d)
    # closed form: multiset difference, one copy per delete key
    vals, cnt = np.unique(fps, return_counts=True)
    dv, dc = np.unique(orc.fingerprints(d, q, r), return_counts=True)
    dmap = dict(zip(dv.tolist(), dc.tolist()))
    rest = np.repeat(vals, [max(0, c - dmap.get(v, 0)) for v, c in zip(vals.tolist(), cnt.tolist())])
    st, want = orc.sqf_canon(rest, q, r, False)
    assert (t == want).all()
    assert removed == len(fps) - len(rest)


def test_sqf_extract_roundtrip_and_capacity(orc):
    q, r = 12, 13
    fps = orc.fingerprints(orc.mt19937(8, 3500), q, r)
    st, can = orc.sqf_canon(fps, q, r, False)
    assert (orc.sqf_extract(can, q, r) == np.sort(fps)).all()
    # more entries than slots -> capacity error (the reference writes out of bounds)
    over = np.arange(orc.sqf_num_slots(q) + 1, dtype=np.uint64) & np.uint64((1 << (q + r)) - 1)
    assert orc.sqf_canon(over, q, r, False)[0] == orc.ERR_CAPACITY


@pytest.mark.parametrize("name", ["rsqf_r5_q10_a0.6.npz", "rsqf_r5_q12_a0.9.npz"])
def test_rsqf_oracle_matches_reference_cpu(orc, name):
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    # faithful insert restatement reproduces the reference's holey layout and return slots
    t = np.zeros(orc.rsqf_table_bytes(q, r), np.uint8)
    fails, ret = orc.rsqf_insert_keys(t, q, r, g["keys"])
    assert (t == g["table_random"]).all()
    assert (ret == g["insert_ret"]).all()
    # faithful lookup on the holey table
    assert (orc.rsqf_lookup_keys(g["table_random"], q, r, g["lookup_keys"]) == g["lookup_random"]).all()
    # canonical closed form == ascending-order reference inserts
    fps = orc.fingerprints(g["keys"], q, r)
    st, can = orc.rsqf_canon(fps, q, r)
    assert (can == g["table_ascending"]).all()
    assert (orc.rsqf_lookup_keys(can, q, r, g["lookup_keys"]) == g["lookup_ascending"]).all()
    assert (orc.rsqf_extract(can, q, r) == np.sort(fps)).all()


def test_rsqf_canonical_r8_invariants(orc):
    # r = 8 has no reference build here (RBITS is fixed in the reference header);
    # pin the r-generic code by invariants: every key found at its canonical slot,
    # extract round trip, and lookup == brute-force "last slot with that fingerprint".
    q, r = 12, 8
    keys = orc.mt19937(21, orc.key_count(0.9, q))
    fps = orc.fingerprints(keys, q, r)
    st, can = orc.rsqf_canon(fps, q, r)
    ext = orc.rsqf_extract(can, q, r)
    assert (ext == np.sort(fps)).all()
    s = np.sort(fps)
    loc = np.zeros(len(s), np.int64)
    fq = (s >> np.uint64(r)).astype(np.int64)
    for i in range(len(s)):
        loc[i] = fq[i] if i == 0 else max(fq[i], loc[i - 1] + 1)
    last = {int(v): int(l) for v, l in zip(s, loc)}
    out = orc.rsqf_lookup_keys(can, q, r, keys)
    assert all(int(o) == last[int(f)] for o, f in zip(out, fps))


def test_fpr_band(orc):
    q, r = 16, 5
    n = orc.key_count(0.8, q)
    keys = orc.mt19937(1, n)
    st, can = orc.sqf_canon(orc.fingerprints(keys, q, r), q, r, True)
    fresh = orc.mt19937(99, 200000)
    fresh = fresh[~np.isin(fresh, keys)]
    hits = (orc.sqf_lookup_keys(can, q, r, fresh) != orc.NOT_FOUND).sum()
    p = n / 2.0 ** (q + r)
    mu, sd = len(fresh) * p, np.sqrt(len(fresh) * p * (1 - p))
    assert abs(hits - mu) < 5 * sd
