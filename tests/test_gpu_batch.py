"""Batched SQFs on the GPU (gqf_sqf_batch_*): many filters of uneven size per
launch must give exactly the per-filter results -- every table equal to
gqf_sqf_bulk_build of the same keys and to the oracle's canonical layout,
every lookup equal to the per-filter unsorted lookup and the oracle -- through
the captured graphs, replayed several times between eager work, and through
the same kernels launched eagerly."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from gpu_helpers import dev, require_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def zipf_keys(rng, n, s=1.1, universe=1 << 28):
    """Zipf(s) ranks over [1, universe] (rejection of the unbounded sampler), as keys rank - 1."""
    out = np.empty(0, np.uint64)
    while len(out) < n:
        z = rng.zipf(s, 2 * (n - len(out)) + 16).astype(np.uint64)
        out = np.concatenate([out, z[z <= universe]])
    return (out[:n] - 1).astype(np.uint32)


def make_batch(F, qlo, qhi, r, seed, fill=(0.3, 0.9), probes=(4, 10)):
    rng = np.random.default_rng(seed)
    qs = [int(x) for x in rng.integers(qlo, qhi + 1, F)]
    keys = [zipf_keys(rng, O.key_count(float(rng.uniform(*fill)), q)) for q in qs]
    probe = []
    for k in keys:
        m = int(2 ** rng.uniform(*probes))
        hit = rng.choice(k, m // 2) if len(k) else np.zeros(0, np.uint32)
        probe.append(np.concatenate([hit, rng.integers(0, 2 ** 32, m - len(hit), dtype=np.uint32)]))
    return qs, keys, probe


def offsets(parts):
    return np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)


def flat(parts):
    return np.concatenate(parts).astype(np.uint32) if parts else np.zeros(0, np.uint32)


def per_filter(qs, r, keys, probe, no_dup):
    """Per-filter calls on fresh (empty) filters: tables, lookups, statuses.
    Dup mode with Zipf keys stores one key ~12% of the time; a run that
    spills past the last slot is a capacity error (table left empty)."""
    tables, looks, status = [], [], []
    for q, k, p in zip(qs, keys, probe):
        f = G.SQF(q, r)
        try:
            f.bulk_build(dev(k), n=len(k), NoDuplicates=no_dup)
            status.append(G.OK)
        except G.CapacityError:
            status.append(G.ERR_CAPACITY)
        tables.append(f.table_bytes())
        out = G.DeviceArray.empty(max(len(p), 1))
        if len(p):
            f.lookup(dev(p), out, n=len(p), sorted=False)
        looks.append(out.to_host()[: len(p)])
        f.destroy()
    return tables, looks, status


def check_batch(qs, r, keys, probe, no_dup, filters, out, ref_tables, ref_looks, ref_status, oracle_every=1):
    po = offsets(probe)
    o = out.to_host()
    for i, (q, k, p) in enumerate(zip(qs, keys, probe)):
        t = filters[i].table_bytes()
        assert (t == ref_tables[i]).all(), ("table", i)
        got = o[int(po[i]):int(po[i + 1])]
        assert (got == ref_looks[i]).all(), ("lookup", i)
        if i % oracle_every == 0:
            st, want = O.sqf_canon(O.fingerprints(k, q, r), q, r, no_dup)
            assert (st == O.OK) == (ref_status[i] == G.OK), ("oracle status", i)
            if st == O.OK:
                assert (t == want).all(), ("oracle table", i)
                assert (got == O.sqf_lookup_keys(want, q, r, p)).all(), ("oracle lookup", i)


@pytest.mark.parametrize("no_dup", [False, True])
def test_batch_equals_per_filter_calls(no_dup):
    r = 5
    qs, keys, probe = make_batch(48, 8, 15, r, seed=7)
    say("sqf batch", len(qs), sum(map(len, keys)), "no_dup", no_dup)
    filters = [G.SQF(q, r) for q in qs]
    dk, dp = dev(flat(keys)), dev(flat(probe))
    out = G.DeviceArray.empty(max(len(flat(probe)), 1))
    b = G.SQFBatch(filters, dk, offsets(keys), dp, offsets(probe), out, no_dup=no_dup)
    info = b.info()
    say("batch plan", info)
    assert info["build_graph_nodes"] > 0 and info["lookup_graph_nodes"] == 3  # index count + scan, lookups
    ref_tables, ref_looks, ref_status = per_filter(qs, r, keys, probe, no_dup)
    # graph replays with eager work (per-filter calls, resets) in between; every replay checked
    for rep in range(3):
        for f in filters:
            f.clear()
        out.fill(0)
        ms, st = b.build(use_graph=True)
        assert st == ref_status
        b.lookup(use_graph=True)
        check_batch(qs, r, keys, probe, no_dup, filters, out, ref_tables, ref_looks, ref_status,
                    oracle_every=5 if rep else 1)
        per_filter(qs[:2], r, keys[:2], probe[:2], no_dup)  # eager work between replays
    for f in filters:
        f.clear()
    out.fill(0)
    assert b.build(use_graph=False)[1] == ref_status
    b.lookup(use_graph=False)
    check_batch(qs, r, keys, probe, no_dup, filters, out, ref_tables, ref_looks, ref_status, oracle_every=7)
    b.destroy()


def test_batch_capacity_empty_and_wide_slots():
    """One filter overfull (capacity: its table is left as it was), one with no
    keys (cleared), 16-bit slots (r = 13)."""
    r = 13
    qs, keys, probe = make_batch(12, 6, 12, r, seed=3)
    keys[4] = np.arange(2 * G.SQF.num_slots(qs[4]), dtype=np.uint32)  # more fingerprints than slots
    keys[7] = np.zeros(0, np.uint32)
    probe[9] = np.zeros(0, np.uint32)
    say("sqf batch capacity/empty", qs)
    filters = [G.SQF(q, r) for q in qs]
    before = []
    for i, f in enumerate(filters):
        f.bulk_build(dev(np.arange(10, dtype=np.uint32)), n=10, NoDuplicates=False)
        before.append(f.table_bytes())
    dk, dp = dev(flat(keys)), dev(flat(probe))
    out = G.DeviceArray.empty(len(flat(probe)))
    b = G.SQFBatch(filters, dk, offsets(keys), dp, offsets(probe), out, no_dup=False)
    for f, t in zip(filters, before):
        f.load_table(t)
    ms, st = b.build(use_graph=True)
    # statuses as the oracle has them (Zipf duplicates overflow some small filters too)
    want = [O.sqf_canon(O.fingerprints(k, q, r), q, r, False) for q, k in zip(qs, keys)]
    assert st == [G.OK if w[0] == O.OK else G.ERR_CAPACITY for w in want]
    assert st[4] == G.ERR_CAPACITY and st[7] == G.OK
    assert (filters[7].table_bytes() == 0).all()
    with pytest.raises(G.CapacityError):
        G.check(G.lib().gqf_sqf_batch_build(b.h, 1, None, None))
    b.lookup(use_graph=True)
    o = out.to_host()
    po = offsets(probe)
    for i, (q, k, p) in enumerate(zip(qs, keys, probe)):
        t = filters[i].table_bytes()
        if st[i] != G.OK:
            assert (t == before[i]).all(), i  # left as it was
            continue
        assert (t == want[i][1]).all(), i
        assert (o[int(po[i]):int(po[i + 1])] == O.sqf_lookup_keys(want[i][1], q, r, p)).all(), i
    b.destroy()


def _hot_batch(F, q, r, seed):
    """Per filter: uniform keys at fill 0.45 plus one hot key repeated 0.3 S
    times (Dup mode), the hot quotient in the first half of the table so the
    run fits.  Probes: the hot key, stored keys whose quotients lie in the
    run's shadow (their walk skips the whole hot run), other stored keys and
    misses."""
    rng = np.random.default_rng(seed)
    S = G.SQF.num_slots(q)
    keys, probe = [], []
    for _ in range(F):
        base = rng.integers(0, 2 ** 32, int(0.45 * S), dtype=np.uint32)
        cand = rng.integers(0, 2 ** 32, 64, dtype=np.uint32)
        cq = (O.fingerprints(cand, q, r).astype(np.uint64) >> r).astype(np.int64)
        hot = cand[np.argmax(cq < (1 << (q - 1)))]
        hq = int(O.fingerprints(np.array([hot], np.uint32), q, r)[0]) >> r
        k = np.concatenate([base, np.full(int(0.3 * S), hot, np.uint32)])
        bq = (O.fingerprints(base, q, r).astype(np.uint64) >> r).astype(np.int64)
        shadow = base[(bq > hq) & (bq < hq + int(0.3 * S) + 64)]
        probe.append(np.concatenate([np.full(8, hot, np.uint32), shadow, rng.choice(base, 200),
                                     rng.integers(0, 2 ** 32, 200, dtype=np.uint32)]))
        keys.append(k)
    return [q] * F, keys, probe


@pytest.mark.parametrize("q,r", [(14, 5), (13, 13), (11, 21)])
def test_batch_long_clusters_lookup_through_the_index(q, r):
    """Lookups whose walk would cross a run of thousands of slots finish
    through the batch's window index; the results equal the per-filter calls
    (plain walk) and the oracle."""
    qs, keys, probe = _hot_batch(6, q, r, seed=11 + q)
    say("sqf batch long clusters", q, r, [len(p) for p in probe])
    filters = [G.SQF(q, r) for q in qs]
    dk, dp = dev(flat(keys)), dev(flat(probe))
    out = G.DeviceArray.empty(len(flat(probe)))
    b = G.SQFBatch(filters, dk, offsets(keys), dp, offsets(probe), out, no_dup=False)
    ref_tables, ref_looks, ref_status = per_filter(qs, r, keys, probe, False)
    assert all(s == G.OK for s in ref_status)
    for use_graph in (True, False):
        out.fill(0)
        ms, st = b.build(use_graph=use_graph)
        assert st == ref_status
        b.lookup(use_graph=use_graph)
        check_batch(qs, r, keys, probe, False, filters, out, ref_tables, ref_looks, ref_status)
    b.destroy()


def test_batch_lookup_sees_tables_changed_after_the_build():
    """A filter changed by a plain call after the batch build (here: a hot key
    inserted into a Dup filter, making a run of thousands of slots): the
    batch lookup indexes the tables as they are, so its results equal the
    per-filter walk on the changed table."""
    q, r = 14, 5
    qs, keys, probe = _hot_batch(4, q, r, seed=5)
    say("sqf batch lookup after a plain change", q, r)
    filters = [G.SQF(q, r) for q in qs]
    dk, dp = dev(flat(keys)), dev(flat(probe))
    out = G.DeviceArray.empty(len(flat(probe)))
    b = G.SQFBatch(filters, dk, offsets(keys), dp, offsets(probe), out, no_dup=False)
    assert b.build(use_graph=True)[1] == [G.OK] * 4
    # filter 2: rebuilt from its keys plus a second hot key (Dup build)
    extra = np.concatenate([keys[2][: len(keys[2]) * 2 // 3], np.full(int(0.2 * G.SQF.num_slots(q)), 424242, np.uint32)])
    st, want = O.sqf_canon(O.fingerprints(extra, q, r), q, r, False)
    assert st == O.OK
    filters[2].load_table(want)
    probe2 = probe[2]
    for use_graph in (True, False):
        out.fill(0)
        b.lookup(use_graph=use_graph)
        po = offsets(probe)
        got = out.to_host()[int(po[2]):int(po[3])]
        assert (got == O.sqf_lookup_keys(want, q, r, probe2)).all()
    b.destroy()


def test_batch_at_the_benchmark_shape():
    """160 filters with q drawn from 12..22 (the batched benchmark's range,
    seed 7), r = 5, fill 0.3..0.9, Zipf(1.1) keys in Dup mode (capacity errors
    included), 2^6..2^12 lookups each: one graph build and one graph lookup
    launch against the per-filter calls for every filter, and the oracle for
    every eighth filter."""
    r = 5
    qs, keys, probe = make_batch(160, 12, 22, r, seed=7, probes=(6, 12))
    say("sqf batch at benchmark shape", len(qs), sum(map(len, keys)))
    filters = [G.SQF(q, r) for q in qs]
    dk, dp = dev(flat(keys)), dev(flat(probe))
    out = G.DeviceArray.empty(len(flat(probe)))
    b = G.SQFBatch(filters, dk, offsets(keys), dp, offsets(probe), out, no_dup=False)
    ref_tables, ref_looks, ref_status = per_filter(qs, r, keys, probe, False)
    for f in filters:
        f.clear()
    out.fill(0)
    ms, st = b.build(use_graph=True)
    assert st == ref_status
    assert ref_status.count(G.OK) > 20 and ref_status.count(G.ERR_CAPACITY) > 20
    b.lookup(use_graph=True)
    check_batch(qs, r, keys, probe, False, filters, out, ref_tables, ref_looks, ref_status, oracle_every=8)
    b.destroy()
