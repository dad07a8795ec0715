"""The reference-named C++ API (include/gqf_compat.hpp) driven by a program
written like the reference drivers (tests/shim/compat_driver.cpp)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "shim", "compat_driver.cpp")
BIN = os.path.join(ROOT, "tests", "shim", "compat_driver")


def compile_driver(out, rbits=8):
    """g++ only: the shim and the driver are plain C++ over the C ABI."""
    libdir = os.path.dirname(G.LIB_PATH)
    # the in-tree binary finds the library relative to itself (the tree moves)
    rpath = "$ORIGIN/../../gpuquotientfilters_amd" if os.path.abspath(out) == BIN else libdir
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-DRBITS={rbits}", "-I", os.path.join(ROOT, "include"),
           SRC, "-L", libdir, "-lgqf", f"-Wl,-rpath,{rpath}", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def fnv1a(b):
    h = 1469598103934665603
    for x in np.frombuffer(np.ascontiguousarray(b).tobytes(), np.uint8).tolist():
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def parse(out):
    return [line.split() for line in out.strip().splitlines()]


def test_shim_compiles_and_host_values_match():
    if not os.path.exists(G.LIB_PATH):
        pytest.skip("libgqf.so not built")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "drv")
        compile_driver(exe)
        out = subprocess.run([exe, "host"], check=True, capture_output=True, text=True).stdout
    rows = parse(out)
    for row in rows:
        if row[0] == "sizeof_quotient_filter" or row[0] == "offset_table":
            assert row[1] == row[2], row  # same layout as the C ABI struct
        elif row[0] == "sizeof_cqf_gpu_block":
            assert int(row[1]) == 1 + 8 + 8 + 8 * 8  # packed, RBITS = 8
        elif row[0] == "slots":
            assert int(row[3]) == O.sqf_table_bytes(int(row[1]), int(row[2]))
        elif row[0] == "blocks":
            assert int(row[2]) == O.rsqf_num_blocks(int(row[1]))
        elif row[0] == "ap":
            k = np.array([int(row[1])], np.uint32)
            assert int(row[2]) == int(O.fingerprints(k, 26, 5)[0])  # AP mod 2^31
            assert int(row[3]) == int(O.fingerprints(k, 18, 5)[0])  # AP mod 2^23


@pytest.mark.gpu
@pytest.mark.parametrize("q,r", [(16, 5), (14, 13)])
def test_shim_sqf_driver_on_gpu(q, r):
    assert os.path.exists(BIN), "tests/shim/compat_driver is built by __graft_entry__.build()"
    n = O.key_count(0.7, q)
    print("[gpu-case] compat driver sqf", q, r, n, flush=True)
    out = subprocess.run([BIN, "sqf", str(q), str(r), str(n), "1"], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    got = {row[0]: row[1:] for row in parse(out)}
    keys = O.mt19937(1, n + n // 4)
    st, t = O.sqf_canon(O.fingerprints(keys[:n], q, r), q, r, True)
    assert got["build"][0] == fnv1a(t)
    assert got["lookup"] == [fnv1a(O.sqf_lookup_keys(t, q, r, keys)), "1"]
    assert O.sqf_insert_keys(t, q, r, keys[n:n + n // 4]) == O.OK
    assert got["insert"][0] == fnv1a(t)
    O.sqf_delete_keys(t, q, r, keys[: n // 8])
    assert got["delete"][0] == fnv1a(t)


@pytest.mark.gpu
def test_shim_rsqf_driver_on_gpu():
    assert os.path.exists(BIN), "tests/shim/compat_driver is built by __graft_entry__.build()"
    q, r = 14, 8
    n = O.key_count(0.8, q)
    print("[gpu-case] compat driver rsqf", q, n, flush=True)
    out = subprocess.run([BIN, "rsqf", str(q), str(n), "1"], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    got = {row[0]: row[1:] for row in parse(out)}
    keys = O.mt19937(1, n)
    st, t = O.rsqf_canon(O.fingerprints(keys, q, r), q, r)
    assert got["insert"][0] == fnv1a(t)
    assert got["lookup"] == [fnv1a(O.rsqf_lookup_keys(t, q, r, keys)), "1"]
