"""CPU check of the GPU torn solve (kronspice_amd/csrc/kh_tear.h): the header
compiled for the host and run by tests/tear_host.cpp against the oracle's
restatement of kron.JacobianTearing (oracle/tear.hpp) on real and complex
systems, B per batch, over call sequences that reuse, refactor, move within
the reuse tolerance, make a piece singular and plant a NaN. Bar: the same
plan, the same decisions (torn, untorn, non-finite), bit-identical torn
solutions and equal counters."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tear_host(tmp_path):
    exe = str(tmp_path / "tear_host")
    csrc = os.path.join(ROOT, "kronspice_amd", "csrc")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-DKH_HD=", "-I", csrc,
           os.path.join(ROOT, "tests", "tear_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
