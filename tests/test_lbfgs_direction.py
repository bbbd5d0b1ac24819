"""The L-BFGS direction solve of the device-resident optimizer on the host
(plastimatch_amd/csrc/lbfgs_logic.h, LbMachine::solve_direction, compiled
here with hipcc as host code by tests/lbfgs_direction_check.cpp): for every
history length m = 1 .. 16, over 3m + 4 iterations of random pairs (the
history wrapping around), the direction the compact form gives equals
liblbfgs's two-loop recursion (libs/liblbfgs-1.9/lib/lbfgs.c:555-609) on
the same vectors to rounding, and the register form of the solve (b <= 6)
gives the bits of the runtime-length form."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not found")
def test_direction_solve_matches_two_loop_recursion(tmp_path):
    exe = tmp_path / "lbfgs_direction_check"
    hipcc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")
    cmd = [hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "plastimatch_amd", "csrc"),
           os.path.join(ROOT, "tests", "lbfgs_direction_check.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.splitlines() if l.startswith("m ")]
    assert [int(l[1]) for l in lines] == list(range(1, 17))
    for l in lines:
        m, maxrel, bitdiff = int(l[1]), float(l[3]), int(l[5])
        assert maxrel < 1e-10, (m, maxrel)
        assert bitdiff == 0, (m, bitdiff)
