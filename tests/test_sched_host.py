"""CPU check of the compiled-schedule LU (kronspice_amd/csrc/kh_sched.h): the
header compiled for the host and run by tests/sched_host.cpp against the
per-lane LU of kh_lu.h on real and complex systems. Bar: every system the
lockstep path accepts has the per-lane LU's pivot rows and bit-identical L, U
and solution; every system it refuses has other pivot rows, or fails (singular,
non-finite) in the per-lane LU too."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_sched_lu_host(tmp_path):
    exe = str(tmp_path / "sched_host")
    csrc = os.path.join(ROOT, "kronspice_amd", "csrc")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-DKH_HD=", "-I", csrc,
           os.path.join(ROOT, "tests", "sched_host.cpp"), os.path.join(csrc, "kh_symbolic.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
