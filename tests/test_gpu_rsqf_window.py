"""RSQF windowed insert (in place, one 4096-slot window per workgroup) against
the decode-and-rebuild insert and the canonical layout of the oracle.

Both paths must leave identical tables and statuses after every batch; the
windowed path must actually run (its kernels launched, no rebuild) wherever
no window exceeds its limits, and fall back to the rebuild (with the same
end state) where one does."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from gpu_helpers import dev, fresh_keys, require_gpu, rsqf_lookup, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def canon(keys, q, r):
    st, t = O.rsqf_canon(O.fingerprints(keys, q, r), q, r)
    assert st == O.OK
    return t


def _kernels(fn):
    G.synchronize()
    G.profile_reset()
    G.profile_enable(True)
    try:
        fn()
        G.synchronize()
    finally:
        G.profile_enable(False)
    names = set(G.profile_stats())
    G.profile_reset()
    return names


def _run(path, q, r, base, batches):
    """Tables, statuses and the kernels each batch launched, on one path."""
    G.set_rsqf_insert_path(path)
    try:
        f = G.RSQF(q, r)
        if len(base):
            f.bulk_build(dev(base))
        tables, status, kern = [], [], []
        for b in batches:
            res = {}

            def ins():
                try:
                    f.insertGPU(dev(b))
                    res["st"] = G.OK
                except G.CapacityError:
                    res["st"] = G.ERR_CAPACITY

            kern.append(_kernels(ins))
            status.append(res["st"])
            tables.append(f.dump())
        f.destroy()
        return tables, status, kern
    finally:
        G.set_rsqf_insert_path(G.INSERT_AUTO)


def _both(q, r, base, batches, expect_window=True):
    tw, sw, kw = _run(G.INSERT_LOCAL, q, r, base, batches)
    tr, sr, _ = _run(G.INSERT_REBUILD, q, r, base, batches)
    assert sw == sr
    for i, (a, b) in enumerate(zip(tw, tr)):
        assert (a == b).all(), ("batch", i, np.flatnonzero(a != b)[:8])
    if expect_window is not None:
        for i, (k, st) in enumerate(zip(kw, sw)):
            if st != G.OK:
                continue
            used = any("k_wm_write" in n for n in k)
            rebuilt = any("k_rr_write" in n for n in k)
            # a fallback launches the windowed kernels (they find the flag and
            # write nothing), then the rebuild
            if expect_window:
                assert used and not rebuilt, ("batch", i, sorted(k))
            else:
                assert rebuilt, ("batch", i, sorted(k))
    return tw, sw


@pytest.mark.parametrize("q,r,alpha", [(12, 5, 0.9), (14, 8, 0.9), (16, 8, 0.95), (17, 16, 0.85), (18, 13, 0.9),
                                       (20, 8, 0.8), (22, 11, 0.9), (24, 10, 0.7)])
def test_window_insert_equals_rebuild_and_canonical(q, r, alpha):
    """Batches from 1 key to 1/8 of the table, 32- and 64-bit fingerprints
    (q + r up to 34), up to fill 0.95."""
    say("rsqf window insert", q, r, alpha)
    n = O.key_count(alpha, q)
    keys = O.mt19937(71, n)
    start = O.key_count(max(alpha - 0.25, 0.05), q)
    cuts = [start, start + 1, start + 17, start + 1000]
    step = max(1, (n - cuts[-1]) // 3)
    cuts += list(range(cuts[-1] + step, n, step)) + [n]
    cuts = sorted(set(c for c in cuts if c <= n))
    batches = [keys[a:b] for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
    tables, status = _both(q, r, keys[:start], batches)
    assert all(s == G.OK for s in status)
    assert (tables[-1] == canon(keys, q, r)).all()


def test_window_insert_tiny_tables_and_table_ends():
    """One window (q = 6..12), keys whose quotients sit in the first and last
    blocks (runs past the last quotient into the spare block), an empty
    table."""
    for q, r in [(6, 8), (7, 5), (8, 16), (12, 8)]:
        say("rsqf window tiny", q, r)
        cand = O.mt19937(72, 1 << 16)
        fq = O.fingerprints(cand, q, r) >> np.uint64(r)
        lo = cand[fq < 3][:6]
        hi = cand[fq >= (1 << q) - 3][:8]
        mid = O.mt19937(73, O.key_count(0.4, q))
        batches = [lo, hi, mid, np.concatenate([hi[:4], lo[:4]])]
        tables, status = _both(q, r, np.zeros(0, np.uint32), batches)
        allk = np.concatenate(batches)
        assert all(s == G.OK for s in status)
        assert (tables[-1] == canon(allk, q, r)).all()


def test_window_insert_duplicates_and_spill_across_windows():
    """Copies of one key: a run that pushes its window's entries a few hundred
    slots into the next window (blocks handed over both ways), and repeated
    batches of the same keys."""
    q, r = 16, 8
    say("rsqf window duplicates", q, r)
    base = O.mt19937(74, O.key_count(0.7, q))
    cand = O.mt19937(75, 4096)
    fq = O.fingerprints(cand, q, r) >> np.uint64(r)
    # a key whose quotient lies 100 slots before a window boundary
    hot = cand[((fq % 4096) > 3900) & ((fq % 4096) < 4000)][0]
    batches = [np.full(150, hot, np.uint32), base[:500], np.full(200, hot, np.uint32), base[:500]]
    tables, status = _both(q, r, base, batches)
    allk = np.concatenate([base] + batches)
    assert (tables[-1] == canon(allk, q, r)).all()


def test_window_insert_falls_back_where_a_window_overflows():
    """9000 copies of one key in one batch: more merged entries than a window
    holds, and a run across a whole window -- the call falls back to the
    rebuild, with the same table; so do later batches on that table."""
    q, r = 15, 8
    say("rsqf window fallback", q, r)
    base = O.mt19937(76, O.key_count(0.4, q))
    cand = O.mt19937(77, 1000)
    hot = cand[(O.fingerprints(cand, q, r) >> np.uint64(r)) < (1 << (q - 2))][0]
    big = np.full(9000, hot, np.uint32)
    tables, status = _both(q, r, base, [big], expect_window=False)
    assert status == [G.OK]
    assert (tables[-1] == canon(np.concatenate([base, big]), q, r)).all()
    extra = fresh_keys(78, 300, base)
    _both(q, r, np.concatenate([base, big]), [extra], expect_window=None)


def test_window_insert_capacity_leaves_table_unchanged():
    q, r = 10, 8
    say("rsqf window capacity", q, r)
    cap = O.rsqf_num_blocks(q) * 64
    base = O.mt19937(79, O.key_count(0.9, q))
    over = O.mt19937(80, cap)
    tables, status = _both(q, r, base, [over, base[:10]], expect_window=None)
    assert status == [G.ERR_CAPACITY, G.OK]
    assert (tables[0] == canon(base, q, r)).all()


def test_window_insert_output_slots():
    """d_out of the windowed path: each key's slot in the final table."""
    q, r = 18, 8
    keys = O.mt19937(81, O.key_count(0.8, q))
    m = len(keys) - 5000
    G.set_rsqf_insert_path(G.INSERT_LOCAL)
    try:
        f = G.RSQF(q, r)
        f.bulk_build(dev(keys[:m]))
        out = G.DeviceArray.empty(5000)
        f.insertGPU(dev(keys[m:]), out)
    finally:
        G.set_rsqf_insert_path(G.INSERT_AUTO)
    t = canon(keys, q, r)
    assert (f.dump() == t).all()
    assert (out.to_host() == O.rsqf_lookup_keys(t, q, r, keys[m:])).all()
    assert (rsqf_lookup(f, keys[:1000]) != G.NOT_FOUND).all()


def test_window_insert_on_reference_tables():
    """Tables written by the reference's own insert (history-dependent layouts
    with holes): both paths decode them by the same rule, so the tables after
    an insert are equal."""
    from conftest import golden
    for name in ["rsqf_r5_q10_a0.6.npz", "rsqf_r5_q12_a0.9.npz"]:
        g = golden(name)
        q, r = int(g["q"]), int(g["r"])
        for tab in ("table_random", "table_ascending"):
            say("rsqf window on reference table", name, tab)
            out = []
            for path in (G.INSERT_LOCAL, G.INSERT_REBUILD):
                G.set_rsqf_insert_path(path)
                try:
                    f = G.RSQF(q, r)
                    f.load(g[tab])
                    f.insertGPU(dev(O.mt19937(82, 40)))
                    out.append(f.dump())
                finally:
                    G.set_rsqf_insert_path(G.INSERT_AUTO)
            assert (out[0] == out[1]).all(), np.flatnonzero(out[0] != out[1])[:8]


@pytest.mark.parametrize("q,r,alpha", [(20, 8, 0.5), (21, 13, 0.85), (19, 5, 0.95)])
def test_window_insert_small_batches_skip_untouched_windows(q, r, alpha):
    """Batches far smaller than the table leave most windows untouched: those
    keep their bytes (no batch entry, same incoming location, no runs into
    the next window) while their neighbours are rewritten around them --
    including a neighbour that ORs its runs into the untouched window's first
    blocks.  Tables equal the rebuild's and the canonical layout after every
    batch, 1 to 2000 keys."""
    say("rsqf window small batches", q, r, alpha)
    n0 = O.key_count(alpha, q)
    keys = O.mt19937(83, n0 + 4000)
    sizes = [1, 2, 7, 50, 300, 2000]
    cuts = np.cumsum([n0] + sizes)
    batches = [keys[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    tables, status = _both(q, r, keys[:n0], batches)
    assert all(s == G.OK for s in status)
    assert (tables[-1] == canon(keys[:cuts[-1]], q, r)).all()
