"""The oracle built with UndefinedBehaviorSanitizer (host code only, g++
-fsanitize=undefined, no recovery): its table operations -- SQF canonical
build, insert, lookup, delete over 8- and 16-bit slots; RSQF canonical build,
the reference-order insert and lookups, 34-bit fingerprints included -- run
through it in a child process, which any shift, overflow or bounds report
aborts.  The checker has to be free of undefined behaviour before its answers
pin anything."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

SCRIPT = r"""
import numpy as np
import oracle as O
for q, r in ((12, 5), (10, 13)):
    keys = O.mt19937(3, O.key_count(0.85, q))
    st, t = O.sqf_canon(O.fingerprints(keys, q, r), q, r, True)
    assert st == O.OK
    probe = np.concatenate([keys[:500], O.mt19937(4, 500)]).astype(np.uint32)
    O.sqf_lookup_keys(t, q, r, probe)
    O.sqf_delete_keys(t, q, r, keys[:300])
    t2 = np.zeros(O.sqf_table_bytes(q, r), np.uint8)
    O.sqf_insert_keys(t2, q, r, keys[:1000])
for q, r in ((12, 5), (20, 14)):
    keys = O.mt19937(5, O.key_count(0.6 if q == 20 else 0.9, q))
    st, t = O.rsqf_canon(O.fingerprints(keys, q, r), q, r)
    assert st == O.OK
    probe = np.concatenate([keys[:500], O.mt19937(6, 500)]).astype(np.uint32)
    O.rsqf_lookup_keys(t, q, r, probe)
    t2 = np.zeros(O.rsqf_table_bytes(q, r), np.uint8)
    O.rsqf_insert_keys(t2, q, r, keys[:2000])
    O.rsqf_extract(t, q, r)
print("ubsan clean")
"""


def test_oracle_operations_are_ubsan_clean(tmp_path):
    out = str(tmp_path / "liboracle_ubsan.so")
    mk = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ubsan", f"UBSAN_OUT={out}"],
                        capture_output=True, text=True)
    if mk.returncode != 0 and "sanitizer" in mk.stderr and "cannot find" in mk.stderr:
        pytest.skip("libubsan not installed")
    assert mk.returncode == 0, mk.stderr[-2000:]
    env = dict(os.environ, GQF_ORACLE_LIB=out, UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               PYTHONPATH=ROOT)
    p = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, env=env, timeout=300,
                       cwd=ROOT)
    assert p.returncode == 0 and "ubsan clean" in p.stdout, (p.stdout + p.stderr)[-3000:]
    assert "runtime error" not in p.stderr, p.stderr[-3000:]
