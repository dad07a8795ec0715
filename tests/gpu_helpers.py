"""Helpers shared by the GPU parity tests (compare the HIP path, through the
C ABI, against the oracle on the same seeded inputs)."""
import sys

import numpy as np

import gpuquotientfilters_amd as G
import oracle as O


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def require_gpu():
    assert G.device_count() > 0, "no GPU visible: the gpu tests need an MI355X"


def dev(a):
    return G.DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.uint32))


def sqf_lookup(qf, keys, sorted=False):
    d = dev(keys)
    out = G.DeviceArray.empty(len(keys))
    qf.lookup(d, out, sorted=sorted)
    return out.to_host()


def rsqf_lookup(f, keys, sorted=False):
    d = dev(keys)
    out = G.DeviceArray.empty(len(keys))
    f.lookup(d, out, sorted=sorted)
    return out.to_host()


def fresh_keys(seed, n, exclude):
    k = O.mt19937(seed, 2 * n + 1024)
    k = k[~np.isin(k, exclude)]
    return k[:n]


def sqf_canon(keys, q, r, no_dup):
    st, t = O.sqf_canon(O.fingerprints(keys, q, r), q, r, no_dup)
    assert st == O.OK
    return t
