"""The SQF's FNV hash option (FNVhashGPU, quotientFilter.cu:117-132) against
the reference CPU_version directly: the golden fixtures hold the reference
CPU insertItem end states (FNV-hashed, packed words) and its mayContain
returns, written by the compiled reference code.  A GPU build of the same
keys must hold the same elements slot for slot, and its lookups must return
what the reference returned -- no oracle in between, only the two format
decoders."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from conftest import golden
from gpu_helpers import dev, require_gpu, say, sqf_lookup

pytestmark = pytest.mark.gpu

NAMES = ["sqf_fnv_q12_r5_a0.6.npz", "sqf_fnv_q14_r5_a0.9.npz", "sqf_fnv_q12_r13_a0.8.npz"]


@pytest.fixture(autouse=True)
def _fnv():
    require_gpu()
    G.set_sqf_hash(G.HASH_FNV)
    yield
    G.set_sqf_hash(G.HASH_AP)


@pytest.mark.parametrize("name", NAMES)
def test_fnv_tables_equal_reference_cpu(name):
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    keys = g["keys"]
    say("sqf fnv vs reference cpu", name)
    ref = O.sqf_decode_words(g["words"], q, r)
    for method in (G.BUILD_SEGMENTED_LAYOUTS, G.BUILD_PARALLEL_MERGING, G.BUILD_SEQUENTIAL_SHIFTS):
        f = G.SQF(q, r)
        f.bulk_build(dev(keys), NoDuplicates=True, method=method)
        assert (O.sqf_decode(f.table_bytes(), q, r) == ref).all(), method
    # incremental: half built, the rest inserted in two batches
    f = G.SQF(q, r)
    h = len(keys) // 2
    f.bulk_build(dev(keys[:h]), NoDuplicates=True)
    f.insert(dev(keys[h:h + h // 2]))
    f.insert(dev(keys[h + h // 2:]))
    assert (O.sqf_decode(f.table_bytes(), q, r) == ref).all()
    for sorted_ in (False, True):
        assert (sqf_lookup(f, g["lookup_keys"], sorted=sorted_) == g["lookup_ret"]).all(), sorted_


def test_fnv_batch_equals_reference_cpu():
    gs = [golden(n) for n in NAMES[:2]]  # both r = 5
    qs = [int(g["q"]) for g in gs]
    keys = [g["keys"] for g in gs]
    probe = [g["lookup_keys"] for g in gs]
    filters = [G.SQF(q, 5) for q in qs]
    koff = np.array([0, len(keys[0]), len(keys[0]) + len(keys[1])], np.uint64)
    poff = np.array([0, len(probe[0]), len(probe[0]) + len(probe[1])], np.uint64)
    out = G.DeviceArray.empty(int(poff[-1]))
    b = G.SQFBatch(filters, dev(np.concatenate(keys)), koff, dev(np.concatenate(probe)), poff, out, no_dup=True)
    b.build(use_graph=True)
    b.lookup(use_graph=True)
    o = out.to_host()
    for i, g in enumerate(gs):
        assert (O.sqf_decode(filters[i].table_bytes(), qs[i], 5) == O.sqf_decode_words(g["words"], qs[i], 5)).all()
        assert (o[int(poff[i]):int(poff[i + 1])] == g["lookup_ret"]).all()
    b.destroy()
