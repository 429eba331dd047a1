"""Builds libkronhip.so (gfx950 kernels + host symbolic phase + C ABI) and the
kh_bench driver (plain C++ over the C ABI) in-tree."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkronhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # Bit-exact with the reference: every a - b*c rounds twice, as in Go.
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
SOURCES = ["kronhip.hip", "kh_symbolic.cpp"]
BENCH = os.path.join(HERE, "kh_bench")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, LIB)
    # kh_bench: the C++ driver (and one copy kernel), linked against the library next to it.
    tmpb = BENCH + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-x", "hip", os.path.join(HERE, "kh_bench.cpp"), "-x", "none", "-L" + HERE, "-lkronhip",
           "-Wl,-rpath,$ORIGIN", "-o", tmpb]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc (kh_bench) failed:\n" + r.stdout + r.stderr)
    os.replace(tmpb, BENCH)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
