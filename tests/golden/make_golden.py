"""Write the golden fixtures under tests/golden/ from the reference CPU code.

The reference CPU_version sources (quotientFilter.cpp, RSQF.cpp, mt19937ar.cpp)
are compiled where they lie by ``oracle/Makefile`` (``make -C oracle ref``)
into ``oracle/_ref/``; this script drives them through the C entry points in
``oracle/ref_*_wrap.cpp`` and stores inputs and outputs as ``.npz`` data.
Run from the repository root:  ``python tests/golden/make_golden.py``
(``--only sqf_packed`` writes just the packed-width SQF fixtures).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
from oracle import RefRSQF, RefSQF  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def save(name, **arrays):
    path = os.path.join(OUT, name)
    np.savez_compressed(path, **arrays)
    print("wrote", name, {k: v.shape for k, v in arrays.items()})


# SQF widths the reference GPU code does not support ((r+3) % 8 != 0); the
# CPU_version packs any width (quotientFilter.cpp:114-130)
SQF_PACKED = [(12, 6, 0.7), (12, 7, 0.75), (12, 8, 0.8), (11, 9, 0.8), (11, 10, 0.85), (10, 11, 0.85), (10, 12, 0.9)]


def sqf_fixture(q, r, alpha):
    """CPU_version SQF: insertItem loop end state + mayContain (quotientFilter.cpp:202-325)."""
    n = O.key_count(alpha, q)
    ks = RefSQF.mt(1, n)
    ref = RefSQF(q, r)
    ins = ref.insert(ks)
    fresh = RefSQF.mt(101, n)
    lk = np.concatenate([ks, fresh])
    save(f"sqf_fnv_q{q}_r{r}_a{alpha}.npz", q=np.uint32(q), r=np.uint32(r), keys=ks,
         insert_ret=ins, words=ref.words(), lookup_keys=lk, lookup_ret=ref.lookup(lk).view(np.uint32))


def main():
    O.build()
    assert RefSQF.available() and RefRSQF.available(), "oracle/_ref not built (needs the reference tree)"
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "sqf_packed":
        for q, r, alpha in SQF_PACKED:
            sqf_fixture(q, r, alpha)
        return

    # MT19937 stream of the reference generator (mt19937ar.cpp:57-70, genrand_int32)
    save("mt19937.npz", seed1=RefSQF.mt(1, 4096), seed7=RefSQF.mt(7, 1024))

    # hashes: Normal_APHash (RSQF.cpp:168-183) and fnv_hash (quotientFilter.cpp:97-112)
    keys = np.concatenate([RefSQF.mt(11, 2048), np.array([0, 1, 0xFF, 0x100, 0xFFFFFFFE, 0xFFFFFFFF], np.uint32)])
    bits = np.array([6, 8, 13, 18, 21, 26, 31], np.uint32)
    ap = np.array([[RefRSQF.lib().refrsqf_ap(int(k), 1 << int(b)) for k in keys] for b in bits], np.uint32)
    fnv = np.array([[RefSQF.fnv(int(k), 1 << int(b)) for k in keys] for b in bits], np.uint32)
    save("hash.npz", keys=keys, bits=bits, ap=ap, fnv=fnv)

    # rankBit / selectBit (RSQF.cpp:185-189, 290-312)
    w = RefSQF.mt(12, 4096).astype(np.uint64)
    words = (w[:2048] << np.uint64(32)) | w[2048:]
    words[::5] &= words[::5] >> np.uint64(7)
    words[::11] = 0
    words[1::13] = np.uint64(0xFFFFFFFFFFFFFFFF)
    ks = np.arange(0, 66, dtype=np.uint32)
    sel = np.array([[RefRSQF.lib().refrsqf_select(int(x), int(k)) for k in ks] for x in words], np.uint32)
    rank = np.array([[RefRSQF.lib().refrsqf_rank(int(x), int(i)) for i in range(64)] for x in words], np.uint32)
    save("rank_select.npz", words=words, ks=ks, select=sel, rank=rank)

    # SQF CPU_version: insertItem loop end state + mayContain (quotientFilter.cpp:202-325)
    for q, r, alpha in [(12, 5, 0.6), (14, 5, 0.9), (12, 13, 0.8)] + SQF_PACKED:
        sqf_fixture(q, r, alpha)

    # RSQF CPU_version: random-order inserts (history-dependent layout) and
    # ascending-fingerprint-order inserts (canonical layout), plus lookups
    r = RefRSQF.lib().refrsqf_rbits()
    for q, alpha in [(10, 0.6), (12, 0.9)]:
        n = O.key_count(alpha, q)
        ks = RefSQF.mt(1, n)
        ref = RefRSQF(q)
        ins = ref.insert(ks)
        fresh = RefSQF.mt(101, n)
        lk = np.concatenate([ks, fresh])
        lr = ref.lookup(lk).view(np.uint32)
        fps = np.array([RefRSQF.lib().refrsqf_ap(int(k), 1 << (q + r)) for k in ks], np.uint32)
        order = np.argsort(fps, kind="stable")
        ref2 = RefRSQF(q)
        ref2.insert(ks[order])
        save(f"rsqf_r{r}_q{q}_a{alpha}.npz", q=np.uint32(q), r=np.uint32(r), keys=ks, insert_ret=ins.view(np.uint32),
             table_random=ref.table(), lookup_keys=lk, lookup_random=lr,
             table_ascending=ref2.table(), lookup_ascending=ref2.lookup(lk).view(np.uint32))


if __name__ == "__main__":
    main()
