"""The serving loop on the GPU (gqf_rsqf_serve): a Poisson request stream
over resident RSQFs, served with coalescing on per-filter streams and again
plainly (one call per request, in arrival order).  Every lookup result and
every final table must be identical between the two, and equal to the oracle
replaying the stream one request at a time."""
import ctypes as C

import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from gpu_helpers import dev, require_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def make_set(qs, r, fill, seed):
    filters, stored = [], []
    for i, q in enumerate(qs):
        keys = O.mt19937(seed + i, O.key_count(fill, q))
        f = G.RSQF(q, r)
        f.bulk_build(dev(keys))
        filters.append(f)
        stored.append(keys)
    return filters, stored


def make_stream(stored, nreq, seed, log2_sizes=(4, 10), mean_gap_us=40.0, big=None):
    """Requests 70/20/10 lookup/insert/delete on uniformly chosen filters.
    Lookups: half stored keys, half uniform; inserts: uniform keys; deletes:
    stored keys, each deleted at most once.  big = (index, op, n) plants one
    oversized request."""
    rng = np.random.default_rng(seed)
    nf = len(stored)
    unused = [rng.permutation(k) for k in stored]
    taken = [0] * nf
    reqs = (G.Request * nreq)()
    hold, plan = [], []
    t = 0.0
    for i in range(nreq):
        t += rng.exponential(mean_gap_us)
        f = int(rng.integers(nf))
        op = int(rng.choice(3, p=[0.7, 0.2, 0.1]))
        n = int(2 ** rng.uniform(*log2_sizes))
        if big is not None and i == big[0]:
            op, n = big[1], big[2]
        if op == G.OP_LOOKUP:
            keys = np.concatenate([rng.choice(stored[f], n // 2), rng.integers(0, 2 ** 32, n - n // 2, dtype=np.uint32)])
        elif op == G.OP_INSERT:
            keys = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
        else:
            n = min(n, len(unused[f]) - taken[f])
            keys = unused[f][taken[f]:taken[f] + n]
            taken[f] += n
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        d = dev(keys)
        out = G.DeviceArray.empty(max(n, 1))
        hold += [d, out]
        reqs[i].filter, reqs[i].op, reqs[i].n = f, op, n
        reqs[i].d_keys, reqs[i].d_out = d.ptr, out.ptr
        reqs[i].arrive_us = t
        plan.append((f, op, keys, out))
    return reqs, plan, hold


def replay_oracle(qs, r, stored, plan, statuses, check_every=3):
    """Sequential oracle: final canonical tables, and lookups (every
    check_every-th) against the canonical table of the state they arrive in."""
    state = [list(O.fingerprints(k, q, r)) for k, q in zip(stored, qs)]
    expect = {}
    for i, (f, op, keys, _) in enumerate(plan):
        if statuses[i] != G.OK:
            continue
        q = qs[f]
        fp = O.fingerprints(keys, q, r)
        if op == G.OP_INSERT:
            state[f].extend(fp.tolist())
        elif op == G.OP_DELETE:
            vals, cnt = np.unique(np.array(state[f], np.uint64), return_counts=True)
            dv, dc = np.unique(fp, return_counts=True)
            dm = dict(zip(dv.tolist(), dc.tolist()))
            state[f] = np.repeat(vals, [max(0, c - dm.get(v, 0)) for v, c in zip(vals.tolist(), cnt.tolist())]).tolist()
        elif i % check_every == 0 and len(keys):
            st, t = O.rsqf_canon(np.array(state[f], np.uint64), q, r)
            assert st == O.OK
            expect[i] = O.rsqf_lookup_keys(t, q, r, keys)
    finals = []
    for f, q in enumerate(qs):
        st, t = O.rsqf_canon(np.array(state[f], np.uint64), q, r)
        assert st == O.OK
        finals.append(t)
    return expect, finals


QS = [13, 14, 15, 16]


def test_served_equals_plain_equals_oracle():
    r = 8
    say("serve coalesced vs plain", QS, r)
    fa, stored = make_set(QS, r, 0.5, seed=31)
    fb, _ = make_set(QS, r, 0.5, seed=31)
    ra, plan_a, hold_a = make_stream(stored, 240, seed=9)
    rb, plan_b, hold_b = make_stream(stored, 240, seed=9)
    sa = G.serve(fa, ra, coalesce=True)
    sb = G.serve(fb, rb, coalesce=False)
    say("serve stats", sa.calls, sb.calls, sa.p99_us, sb.p99_us)
    assert all(x.status == G.OK for x in ra) and all(x.status == G.OK for x in rb)
    assert sa.requests == sb.requests == 240
    assert sb.calls == 240 and sa.calls < 240  # coalescing happened
    for i, ((f, op, keys, oa), (_, _, _, ob)) in enumerate(zip(plan_a, plan_b)):
        if op == G.OP_LOOKUP and len(keys):
            assert (oa.to_host()[: len(keys)] == ob.to_host()[: len(keys)]).all(), i
    expect, finals = replay_oracle(QS, r, stored, plan_a, [x.status for x in ra])
    for i, want in expect.items():
        assert (plan_a[i][3].to_host()[: len(want)] == want).all(), i
    for f in range(len(QS)):
        da = fa[f].dump()
        assert (da == fb[f].dump()).all(), f
        assert (da == finals[f]).all(), f


def test_tight_budget_serialises_and_rejects():
    """A budget just above the tables: calls wait for each other, idle
    workers release their workspaces, and an insert too large to ever fit is
    rejected while the rest of the stream is served correctly."""
    r = 8
    say("serve tight budget", QS, r)
    fa, stored = make_set(QS, r, 0.5, seed=41)
    resident = sum(G.lib().gqf_rsqf_table_bytes(q, r) for q in QS)
    big_n = 1 << 19
    need_big = min(G.lib().gqf_workspace_alloc_bytes(G.lib().gqf_rsqf_workspace_bytes(f.h, G.OP_INSERT, big_n, 0))
                   for f in fa)
    budget = resident + (5 << 20)
    assert need_big > 5 << 20
    ra, plan, hold = make_stream(stored, 160, seed=10, big=(70, G.OP_INSERT, big_n))
    st = G.serve(fa, ra, coalesce=True, hbm_budget=budget)
    statuses = [x.status for x in ra]
    say("tight budget stats", st.calls, st.rejected, st.peak_bytes - resident)
    assert statuses[70] == G.ERR_NOMEM and ra[70].call == -1
    assert st.rejected == 1 and statuses.count(G.OK) == 159
    assert st.peak_bytes <= budget
    expect, finals = replay_oracle(QS, r, stored, plan, statuses)
    for i, want in expect.items():
        assert (plan[i][3].to_host()[: len(want)] == want).all(), i
    for f in range(len(QS)):
        assert (fa[f].dump() == finals[f]).all(), f


def test_wide_fingerprints_in_the_serving_loop():
    """Filters at the serving configuration's sizes: q = 24 and 25 with r = 8
    (32- and 33-bit fingerprints; the second takes the 64-bit decode path)
    beside a small one, under the same mixed stream, coalesced and plain.
    Every lookup and final table agrees between the two; the oracle replays
    the stream: the small filter's lookups as they arrive, each big filter's
    final table and a last lookup batch on it."""
    r = 8
    qs = [16, 24, 25]
    say("serve wide fingerprints", qs, r)
    fa, stored = make_set(qs, r, 0.5, seed=51)
    fb, _ = make_set(qs, r, 0.5, seed=51)
    n = 90
    ra, plan_a, hold_a = make_stream(stored, n, seed=12, log2_sizes=(6, 14), mean_gap_us=60.0)
    rb, plan_b, hold_b = make_stream(stored, n, seed=12, log2_sizes=(6, 14), mean_gap_us=60.0)
    # a last lookup batch on every filter, after everything else has arrived
    rng = np.random.default_rng(13)
    tail = []
    for f in range(len(qs)):
        keys = np.concatenate([rng.choice(stored[f], 3000), rng.integers(0, 2 ** 32, 3000, dtype=np.uint32)])
        tail.append((f, keys.astype(np.uint32)))
    extra_a = (G.Request * (n + len(tail)))()
    extra_b = (G.Request * (n + len(tail)))()
    for src, dst in ((ra, extra_a), (rb, extra_b)):
        C_size = C.sizeof(G.Request)
        C.memmove(C.addressof(dst), C.addressof(src), n * C_size)
    t_end = max(x.arrive_us for x in ra) + 50.0
    outs = []
    for j, (f, keys) in enumerate(tail):
        d = dev(keys)
        oa, ob = G.DeviceArray.empty(len(keys)), G.DeviceArray.empty(len(keys))
        hold_a += [d, oa, ob]
        for req, o in ((extra_a[n + j], oa), (extra_b[n + j], ob)):
            req.filter, req.op, req.n = f, G.OP_LOOKUP, len(keys)
            req.d_keys, req.d_out, req.arrive_us = d.ptr, o.ptr, t_end + j
        outs.append((oa, ob))
    sa = G.serve(fa, extra_a, coalesce=True)
    sb = G.serve(fb, extra_b, coalesce=False)
    assert all(x.status == G.OK for x in extra_a) and all(x.status == G.OK for x in extra_b)
    assert sa.requests == sb.requests == n + len(tail)
    for i, ((f, op, keys, oa), (_, _, _, ob)) in enumerate(zip(plan_a, plan_b)):
        if op == G.OP_LOOKUP and len(keys):
            assert (oa.to_host()[: len(keys)] == ob.to_host()[: len(keys)]).all(), i
    # the oracle: small filter lookups on arrival (the replay checks filter 0
    # only), final tables and the last lookups everywhere
    small_plan = [(f, op, keys, o) if f == 0 or op != G.OP_LOOKUP else (f, op, np.zeros(0, np.uint32), o)
                  for (f, op, keys, o) in plan_a]
    expect, finals = replay_oracle(qs, r, stored, small_plan, [x.status for x in ra[:n]])
    for i, want in expect.items():
        assert (plan_a[i][3].to_host()[: len(want)] == want).all(), i
    for f in range(len(qs)):
        da = fa[f].dump()
        assert (da == fb[f].dump()).all(), f
        assert (da == finals[f]).all(), f
        keys = tail[f][1]
        want = O.rsqf_lookup_keys(finals[f], qs[f], r, keys)
        assert (outs[f][0].to_host() == want).all() and (outs[f][1].to_host() == want).all(), f
