"""Launch planners on the CPU: the radix-sort chunk plan, the rebuild
sub-tile plan and the row-scan pieces, at the sizes that have broken such
planners before (empty input, one key, tile boundaries, the bench sizes, the
per-call maximum).  The planners are host code in the kernel headers; the
check program (tests/plan/plan_check.cpp) includes them and is built for the
host only, so no GPU is needed."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "plan", "plan_check.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_planners_cover_every_size(tmp_path):
    exe = str(tmp_path / "plan_check")
    cmd = [HIPCC, "--cuda-host-only", "-std=c++17", "-O1", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", SRC, "-o", exe]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "plans ok" in r.stdout, r.stdout[-2000:]
