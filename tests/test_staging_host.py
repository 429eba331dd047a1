"""k_persist's LDS staging plan on the CPU (tests/staging_host.cpp compiled
with g++): over groups of many shapes (dense and sparse backends, optional
per-circuit arrays, unallocated LU workspace) and LDS budgets, the staged
arrays of every circuit lie inside its LDS slice without overlapping, at the
offsets the kernel's copy loops use; the arrays left in device memory are
viewed at the circuit's own slices, exactly where the in-place kernel
addresses them; the plan keeps to the budget, stages everything when
everything fits and takes the most used arrays first."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_staging_plan_and_views(tmp_path):
    exe = str(tmp_path / "staging_host")
    csrc = os.path.join(ROOT, "kronspice_amd", "csrc")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-DKH_HD=", "-I",
                        os.path.join(ROOT, "tests", "host_shim"), "-I", csrc,
                        os.path.join(ROOT, "tests", "staging_host.cpp"), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok "), r.stdout + r.stderr
