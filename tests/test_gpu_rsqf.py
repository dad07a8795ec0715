"""RSQF on the GPU through the C ABI.

Lookups must equal the reference lookup on any table -- the golden tables
written by the compiled reference CPU insert (history-dependent layouts with
holes) included.  Build / insert / delete must equal the canonical layout."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from conftest import golden
from gpu_helpers import dev, fresh_keys, require_gpu, rsqf_lookup, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def canon(keys, q, r):
    st, t = O.rsqf_canon(O.fingerprints(keys, q, r), q, r)
    assert st == O.OK
    return t


@pytest.mark.parametrize("name", ["rsqf_r5_q10_a0.6.npz", "rsqf_r5_q12_a0.9.npz"])
@pytest.mark.parametrize("sorted_", [False, True])
def test_lookup_on_reference_tables(name, sorted_):
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    for tab, res in (("table_random", "lookup_random"), ("table_ascending", "lookup_ascending")):
        say("rsqf golden lookup", name, tab, sorted_)
        f = G.RSQF(q, r)
        f.load(g[tab])
        assert (f.dump() == g[tab]).all()
        got = rsqf_lookup(f, g["lookup_keys"], sorted=sorted_)
        assert (got == g[res]).all()


CASES = [(10, 5, 0.6), (12, 5, 0.9), (14, 8, 0.9), (16, 8, 0.95), (18, 16, 0.9), (20, 5, 0.9), (20, 8, 0.8)]


@pytest.mark.parametrize("q,r,alpha", CASES)
def test_build_equals_canonical(q, r, alpha):
    say("rsqf build", q, r, alpha)
    n = O.key_count(alpha, q)
    keys = O.mt19937(1, n)
    keys[n - n // 8:] = keys[: n // 8]  # duplicates are stored
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys))
    want = canon(keys, q, r)
    got = f.dump()
    assert (got == want).all(), np.flatnonzero(got != want)[:8]


@pytest.mark.parametrize("q,r,alpha", [(12, 5, 0.9), (16, 8, 0.9), (18, 16, 0.9), (20, 8, 0.9)])
@pytest.mark.parametrize("sorted_", [False, True])
def test_lookup_matches_oracle(q, r, alpha, sorted_):
    say("rsqf lookup", q, r, alpha, sorted_)
    n = O.key_count(alpha, q)
    keys = O.mt19937(1, n)
    t = canon(keys, q, r)
    f = G.RSQF(q, r)
    f.load(t)
    probe = np.concatenate([keys, fresh_keys(3, n, keys)])
    got = rsqf_lookup(f, probe, sorted=sorted_)
    assert (got == O.rsqf_lookup_keys(t, q, r, probe)).all()
    assert (got[:n] != G.NOT_FOUND).all()


@pytest.mark.parametrize("q,r", [(12, 5), (16, 8), (18, 16)])
def test_insert_batches_equal_canonical(q, r):
    say("rsqf insert", q, r)
    n = O.key_count(0.9, q)
    keys = O.mt19937(2, n)
    start = O.key_count(0.7, q)
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys[:start]))
    step = max(1, 1 << (q - 6))
    for i in range(start, n, step):
        b = keys[i:i + step]
        out = G.DeviceArray.empty(len(b))
        f.insertGPU(dev(b), out)
    want = canon(keys, q, r)
    assert (f.dump() == want).all()
    # the per-key output is the key's slot in the final table
    probe = keys[start:]
    assert (rsqf_lookup(f, probe) == O.rsqf_lookup_keys(want, q, r, probe)).all()


def test_insert_output_is_final_slot():
    q, r = 12, 8
    keys = O.mt19937(8, 3000)
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys[:2000]))
    out = G.DeviceArray.empty(1000)
    f.insertGPU(dev(keys[2000:]), out)
    t = canon(keys, q, r)
    assert (out.to_host() == O.rsqf_lookup_keys(t, q, r, keys[2000:])).all()


@pytest.mark.parametrize("q,r,frac", [(12, 5, 0.1), (16, 8, 0.5), (18, 16, 0.2)])
def test_delete_equals_canonical(q, r, frac):
    say("rsqf delete", q, r, frac)
    n = O.key_count(0.9, q)
    keys = O.mt19937(4, n)
    base = np.concatenate([keys, keys[:100]])
    f = G.RSQF(q, r)
    f.bulk_build(dev(base))
    d = np.concatenate([keys[: int(n * frac)], keys[:50], O.mt19937(44, 300)])
    f.deleteGPU(dev(d))
    fb = O.fingerprints(base, q, r)
    vals, cnt = np.unique(fb, return_counts=True)
    dv, dc = np.unique(O.fingerprints(d, q, r), return_counts=True)
    dmap = dict(zip(dv.tolist(), dc.tolist()))
    rest = np.repeat(vals, [max(0, c - dmap.get(v, 0)) for v, c in zip(vals.tolist(), cnt.tolist())])
    st, want = O.rsqf_canon(rest, q, r)
    assert (f.dump() == want).all()


def test_edge_cases_and_capacity():
    q, r = 8, 8
    f = G.RSQF(q, r)
    f.bulk_build(dev(np.zeros(0, np.uint32)), n=0)
    assert (f.dump() == 0).all()
    k = np.array([0xFFFFFFFF, 0, 7], np.uint32)
    f.insertGPU(dev(k))
    assert (f.dump() == canon(k, q, r)).all()
    before = f.dump()
    cap = O.rsqf_num_blocks(q) * 64
    with pytest.raises(G.CapacityError):
        f.insertGPU(dev(np.arange(cap + 5, dtype=np.uint32)))
    assert (f.dump() == before).all()


def test_full_size_q26_r8_properties():
    """RSQF at q=26, r=8 (34-bit wide fingerprints), fill 0.9: no false
    negatives, sorted == unsorted, FPR band, stored-entry count."""
    q, r = 26, 8
    n = O.key_count(0.9, q)
    say("rsqf full size", q, r, n)
    keys = O.mt19937(1, n)
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys))
    m = 1 << 20
    probe = np.concatenate([keys[: m // 2], fresh_keys(5, m // 2, keys)])
    u = rsqf_lookup(f, probe, sorted=False)
    s = rsqf_lookup(f, probe, sorted=True)
    assert (u == s).all()
    assert (u[: m // 2] != G.NOT_FOUND).all()
    # a quotient filter is exact on fingerprints: hit <=> fingerprint stored
    fk = O.fingerprints(keys, q, r)
    assert ((u != G.NOT_FOUND) == np.isin(O.fingerprints(probe, q, r), fk)).all()
    # the wide fingerprint is uniform: FPR at alpha * 2^-r
    p = n / 2.0 ** (q + r)
    rate = (u[m // 2:] != G.NOT_FOUND).mean()
    assert abs(rate - p) < 6 * np.sqrt(p * (1 - p) / (m // 2))
    t = f.dump()
    off, occ, ends, rem = O.rsqf_fields(t, q, r)
    runs = int(sum(bin(int(x)).count("1") for x in ends[:: 97]))
    occs = int(sum(bin(int(x)).count("1") for x in occ[:: 97]))
    assert runs > 0 and occs > 0
    total_runs = int(np.unpackbits(ends.view(np.uint8)).sum())
    total_occ = int(np.unpackbits(occ.view(np.uint8)).sum())
    assert total_runs == total_occ == len(np.unique(O.fingerprints(keys, q, r) >> np.uint64(r)))


@pytest.mark.parametrize("q,r", [(14, 8), (14, 5), (15, 16)])
def test_long_runs_across_windows(q, r):
    """A key stored 5000 times: its run spans more than one 4096-slot decode
    window and the following blocks' offsets pass 255 (stored mod 256).
    Insert and delete decode such a table through run ranks, not offsets."""
    say("rsqf long runs", q, r)
    # a key whose quotient lies in the first quarter, so the run fits
    cand = O.mt19937(5, 1000)
    hot = cand[(O.fingerprints(cand, q, r) >> np.uint64(r)) < (1 << (q - 2))][0]
    base = np.concatenate([O.mt19937(21, 6000), np.full(5000, hot, np.uint32)])
    f = G.RSQF(q, r)
    f.bulk_build(dev(base))
    t = f.dump()
    assert (t == canon(base, q, r)).all()
    # lookups follow the reference walk, which reads the wrapped offsets: in
    # the blocks behind the long run it can land on the wrong run and miss a
    # stored key.  The GPU answers must be the reference walk's, misses
    # included (the oracle's lookup restates RSQF.cu:320-399 and is pinned to
    # the reference's own tables).
    probe = np.concatenate([O.mt19937(21, 6000), [hot], fresh_keys(23, 3000, base)]).astype(np.uint32)
    want = O.rsqf_lookup_keys(t, q, r, probe)
    assert (rsqf_lookup(f, probe) == want).all()
    assert (rsqf_lookup(f, probe, sorted=True) == want).all()
    extra = np.concatenate([O.mt19937(22, 1500), np.full(300, hot, np.uint32)])
    f.insertGPU(dev(extra))
    allk = np.concatenate([base, extra])
    assert (f.dump() == canon(allk, q, r)).all()
    d = np.full(4000, hot, np.uint32)
    f.deleteGPU(dev(d))
    rest = np.concatenate([O.mt19937(21, 6000), np.full(1300, hot, np.uint32), O.mt19937(22, 1500)])
    assert (f.dump() == canon(rest, q, r)).all()

@pytest.mark.parametrize("q,r,alpha", [(25, 8, 0.85), (22, 13, 0.6)])
def test_wide_insert_delete_equal_canonical(q, r, alpha):
    """Fingerprints wider than 32 bits (q + r = 33, 35): insert batches and a
    delete against the canonical layout."""
    say("rsqf wide insert/delete", q, r, alpha)
    n = O.key_count(alpha, q)
    keys = O.mt19937(6, n)
    start = O.key_count(alpha - 0.1, q)
    f = G.RSQF(q, r)
    f.bulk_build(dev(keys[:start]))
    step = (n - start + 2) // 3
    for i in range(start, n, step):
        f.insertGPU(dev(keys[i:i + step]))
    want = canon(keys, q, r)
    assert (f.dump() == want).all()
    d = keys[: n // 10]
    f.deleteGPU(dev(d))
    assert (f.dump() == canon(keys[n // 10:], q, r)).all()


@pytest.mark.parametrize("q,r,hot", [(14, 8, 200), (13, 5, 200), (20, 14, 200), (14, 8, 5000), (15, 16, 3000)])
def test_count_query(q, r, hot):
    """gqf_rsqf_count == multiplicity of each key's fingerprint in the stored
    multiset (duplicates, a run of `hot` copies of one key, misses; 34-bit
    fingerprints at q=20 r=14).  Counts follow run ranks, not the 8-bit
    offsets: a run of thousands of copies pushes the offsets of the blocks
    behind it past 255, and every count there stays exact."""
    say("rsqf count", q, r, hot)
    n = O.key_count(0.7 if hot <= 200 else 0.3, q)
    keys = O.mt19937(31, n)
    cand = O.mt19937(5, 1000)
    hotk = cand[(O.fingerprints(cand, q, r) >> np.uint64(r)) < (1 << (q - 2))][0]
    stored = np.concatenate([keys, keys[:n // 5], keys[:n // 20], np.full(hot, hotk, np.uint32)])
    f = G.RSQF(q, r)
    f.bulk_build(dev(stored))
    probe = np.concatenate([keys, [hotk], fresh_keys(32, 2000, stored)]).astype(np.uint32)
    out = G.DeviceArray.empty(len(probe))
    f.count(dev(probe), out)
    vals, cnt = np.unique(O.fingerprints(stored, q, r), return_counts=True)
    want = dict(zip(vals.tolist(), cnt.tolist()))
    exp = np.array([want.get(int(x), 0) for x in O.fingerprints(probe, q, r)], np.uint32)
    got = out.to_host()
    assert (got == exp).all(), np.flatnonzero(got != exp)[:8]
    assert got[n] >= hot


@pytest.mark.parametrize("q,r,alpha", [(24, 10, 0.2), (26, 8, 0.05)])
def test_wide_decode_bounds_crossed_by_a_long_run(q, r, alpha):
    """34-bit fingerprints decode into 32-bit low words plus the indices where
    the high word steps up.  6000 copies of a key whose quotient lies just
    below the first step form a run that reaches into later 4096-slot decode
    windows; insert and delete against the canonical layout."""
    say("rsqf wide decode bounds", q, r)
    step = 1 << (32 - r)  # quotient where the high word becomes 1
    cand = O.mt19937(7, 1 << 21)
    fq = O.fingerprints(cand, q, r) >> np.uint64(r)
    hot = cand[(fq >= step - 300) & (fq < step - 100)][0]
    rnd = O.mt19937(8, O.key_count(alpha, q))
    base = np.concatenate([rnd, np.full(6000, hot, np.uint32)])
    f = G.RSQF(q, r)
    f.bulk_build(dev(base))
    extra = np.concatenate([O.mt19937(9, 50000), np.full(500, hot, np.uint32)])
    f.insertGPU(dev(extra))
    assert (f.dump() == canon(np.concatenate([base, extra]), q, r)).all()
    f.deleteGPU(dev(np.concatenate([np.full(3000, hot, np.uint32), rnd[:20000]])))
    rest = np.concatenate([rnd[20000:], extra[:50000], np.full(3500, hot, np.uint32)])
    assert (f.dump() == canon(rest, q, r)).all()


@pytest.mark.parametrize("q,r,fill,extra", [(10, 8, 0.9, 400), (12, 5, 0.95, 300), (18, 16, 0.97, 12000)])
def test_insert_report_places_what_fits(q, r, fill, extra):
    """The reference's insertGPU writes -1 for an item it cannot place and
    keeps the rest (RSQF.cu:745-860).  gqf_rsqf_insert_report: with e = 1 +
    the last location of the whole merge and S slots, the keys with the e - S
    largest fingerprints (all copies of the cut fingerprint) are left out,
    the rest is inserted, and the table is the canonical layout of what was
    placed.  A batch that fits behaves as the plain insert."""
    say("rsqf insert report", q, r, fill, extra)
    base = O.mt19937(91, O.key_count(fill, q))
    batch = np.concatenate([fresh_keys(92, extra, base), base[:20]])  # duplicates of stored keys too
    S = O.rsqf_num_blocks(q) * 64
    f = G.RSQF(q, r)
    f.bulk_build(dev(base))
    out = G.DeviceArray.empty(len(batch))
    _, failed = f.insert_report(dev(batch), out)
    # expected cut, from the canonical layout of everything
    allfp = np.sort(np.concatenate([O.fingerprints(base, q, r), O.fingerprints(batch, q, r)]))
    fq = (allfp >> np.uint64(r)).astype(np.int64)
    i = np.arange(len(fq))
    e = int((np.maximum.accumulate(fq - i) + i)[-1]) + 1
    assert e > S, "the case must overflow"
    bf = O.fingerprints(batch, q, r)
    cut = np.sort(bf)[len(batch) - (e - S)]
    keep = bf < cut
    assert failed == int((~keep).sum()) > 0
    t = canon(np.concatenate([base, batch[keep]]), q, r)
    assert (f.dump() == t).all()
    got = out.to_host().view(np.int32)
    assert (got[~keep] == -1).all()
    assert (got[keep] == O.rsqf_lookup_keys(t, q, r, batch[keep]).view(np.int32)).all()
    # a batch that fits: plain insert, nothing left out
    g2 = G.RSQF(q, r)
    g2.bulk_build(dev(base[: len(base) // 2]))
    out2 = G.DeviceArray.empty(50)
    _, failed2 = g2.insert_report(dev(base[-50:]), out2)
    assert failed2 == 0
    t2 = canon(np.concatenate([base[: len(base) // 2], base[-50:]]), q, r)
    assert (g2.dump() == t2).all()
    assert (out2.to_host() == O.rsqf_lookup_keys(t2, q, r, base[-50:])).all()
