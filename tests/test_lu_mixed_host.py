"""CPU checks of the mixed-precision LU: kronspice_amd/csrc/kh_lu.h compiled
for the host (it is host+device code) and run by tests/lu_mixed_host.cpp —
certified pivot rows equal to the double-precision factorization's on real
MNA systems, and complex lanes in structure-of-arrays layout whose phases
interleave as concurrent wavefronts do (each lane's single-precision data
must stay inside its own entries)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mixed_lu_host_build(tmp_path):
    exe = str(tmp_path / "lu_mixed_host")
    csrc = os.path.join(ROOT, "kronspice_amd", "csrc")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-DKH_HD=", "-I", csrc,
           os.path.join(ROOT, "tests", "lu_mixed_host.cpp"), os.path.join(csrc, "kh_symbolic.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
