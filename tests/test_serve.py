"""The serving scheduler (gqf_serve_run) against a host model of the device:
no GPU.  The model's calls sleep in proportion to their keys and track the
device bytes each worker holds, so the tests can check what the scheduler
promises: per-filter arrival order, same-kind coalescing within max_batch,
the memory budget at every call start, rejection of requests that cannot fit
on their own, and one request per call in plain mode."""
import ctypes as C
import threading
import time

import numpy as np
import pytest

import gpuquotientfilters_amd as G

PER_KEY = 16  # model workspace bytes per key of a call
STAGE = 8     # model gather bytes per key of a coalesced call


def make_requests(nfilters, nreq, seed, log2_sizes=(4, 10), mean_gap_us=40.0):
    rng = np.random.default_rng(seed)
    reqs = (G.Request * nreq)()
    t = 0.0
    for i in range(nreq):
        t += rng.exponential(mean_gap_us)
        reqs[i].filter = int(rng.integers(nfilters))
        reqs[i].op = int(rng.choice(3, p=[0.7, 0.2, 0.1]))
        reqs[i].n = int(2 ** rng.uniform(*log2_sizes))
        reqs[i].d_keys = 0x1000  # the model never dereferences buffers
        reqs[i].d_out = 0x1000
        reqs[i].arrive_us = t
    return reqs


class Model:
    def __init__(self, reqs, nworkers, resident, budget, ns_per_key=100):
        self.base = C.addressof(reqs[0])
        self.ws = [0] * nworkers
        self.resident, self.budget = resident, budget
        self.ns_per_key = ns_per_key
        self.lock = threading.Lock()
        self.calls = []  # (worker, [request indices]) in start order
        self.max_used = 0
        self.violations = 0
        self.ex = G.ServeExecutor(None, G.GROW_FN(self.grow), G.HELD_FN(self.held), G.TRIM_FN(self.trim),
                                  G.RUN_FN(self.run))

    @staticmethod
    def need(n, coalesced):
        return PER_KEY * n + (STAGE * n if coalesced else 0)

    def grow(self, _u, w, f, op, n, coalesced):
        return max(0, self.need(n, coalesced) - self.ws[w])

    def held(self, _u, w):
        return self.ws[w]

    def trim(self, _u, w):
        self.ws[w] = 0

    def run(self, _u, w, batch, nb):
        idx = [(C.addressof(batch[i].contents) - self.base) // C.sizeof(G.Request) for i in range(nb)]
        n = sum(batch[i].contents.n for i in range(nb))
        with self.lock:
            self.ws[w] = max(self.ws[w], self.need(n, nb > 1))
            used = self.resident + sum(self.ws)
            self.max_used = max(self.max_used, used)
            if used > self.budget:
                self.violations += 1
            self.calls.append((w, idx))
        time.sleep(20e-6 + n * self.ns_per_key * 1e-9)
        return G.OK


def serve(reqs, nfilters, coalesce, budget, resident=1 << 20, max_batch=0, ns_per_key=100):
    m = Model(reqs, nfilters if coalesce else 1, resident, budget, ns_per_key)
    cfg = G.ServeConfig(budget, int(coalesce), 0, max_batch)
    st = G.ServeStats()
    status = G.lib().gqf_serve_run(nfilters, resident, reqs, len(reqs), C.byref(cfg), C.byref(m.ex), C.byref(st))
    return status, st, m


def per_filter_order_holds(reqs, m):
    """Every call covers consecutive requests of one filter and kind; per
    filter, calls in call-id order replay the arrival order."""
    by_call = {}
    for w, idx in m.calls:
        assert len({reqs[i].filter for i in idx}) == 1
        assert len({reqs[i].op for i in idx}) == 1
        assert len({reqs[i].call for i in idx}) == 1
        by_call[reqs[idx[0]].call] = idx
    nf = max(r.filter for r in reqs) + 1
    for f in range(nf):
        seq = [i for c in sorted(by_call) for i in by_call[c] if reqs[i].filter == f]
        want = [i for i in range(len(reqs)) if reqs[i].filter == f and reqs[i].status == G.OK]
        assert seq == want
    return by_call


def test_coalescing_keeps_per_filter_order():
    reqs = make_requests(4, 400, seed=1, mean_gap_us=15.0)
    status, st, m = serve(reqs, 4, True, budget=1 << 40)
    assert status == G.OK and m.violations == 0
    assert all(r.status == G.OK for r in reqs)
    by_call = per_filter_order_holds(reqs, m)
    assert st.calls == len(by_call) < len(reqs)  # arrivals outpace service: batches form
    assert max(len(v) for v in by_call.values()) > 1
    assert st.requests == len(reqs) and st.rejected == 0
    assert st.keys == sum(r.n for r in reqs)
    lat = np.array([r.done_us - r.arrive_us for r in reqs])
    assert (lat > 0).all()
    assert st.p50_us <= st.p99_us <= lat.max() + 1e-6
    assert abs(st.p50_us - np.sort(lat)[int(np.ceil(0.5 * len(lat))) - 1]) < 1e-6


def test_max_batch_bounds_coalesced_calls():
    reqs = make_requests(2, 300, seed=2, mean_gap_us=5.0)
    status, st, m = serve(reqs, 2, True, budget=1 << 40, max_batch=3000)
    assert status == G.OK
    per_filter_order_holds(reqs, m)
    for _, idx in m.calls:
        if len(idx) > 1:
            assert sum(reqs[i].n for i in idx) <= 3000


def test_plain_mode_one_request_per_call_in_arrival_order():
    reqs = make_requests(3, 200, seed=3)
    status, st, m = serve(reqs, 3, False, budget=1 << 40)
    assert status == G.OK
    assert [idx for _, idx in m.calls] == [[i] for i in range(len(reqs))]
    assert st.calls == len(reqs)
    assert [r.call for r in reqs] == list(range(len(reqs)))


@pytest.mark.parametrize("coalesce", [True, False])
def test_budget_holds_and_oversized_requests_are_rejected(coalesce):
    reqs = make_requests(4, 300, seed=4, log2_sizes=(4, 11), mean_gap_us=10.0)
    resident = 1 << 20
    room = PER_KEY * 600  # a request of more than 600 keys can never run
    status, st, m = serve(reqs, 4, coalesce, budget=resident + room, resident=resident)
    assert status == G.OK
    assert m.violations == 0 and m.max_used <= resident + room
    assert st.peak_bytes <= resident + room
    rejected = [i for i, r in enumerate(reqs) if r.status == G.ERR_NOMEM]
    assert rejected == [i for i, r in enumerate(reqs) if PER_KEY * r.n > room]
    assert st.rejected == len(rejected) > 0
    assert all(reqs[i].call == -1 for i in rejected)
    per_filter_order_holds(reqs, m)


def test_invalid_streams_are_refused():
    reqs = make_requests(2, 10, seed=5)
    reqs[5].arrive_us = reqs[4].arrive_us - 1.0
    status, _, m = serve(reqs, 2, True, budget=1 << 40)
    assert status == G.ERR_INVALID and not m.calls
    reqs = make_requests(2, 10, seed=5)
    reqs[3].filter = 2
    status, _, m = serve(reqs, 2, True, budget=1 << 40)
    assert status == G.ERR_INVALID and not m.calls
