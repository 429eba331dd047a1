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
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=gfx950",
    # Bit-exact with the reference: every a - b*c rounds twice, as in Go.
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# (source, extra flags, object): the lane-per-circuit persistent kernel once
# per register budget (resident wavefronts per SIMD) and element-kind set, the
# wavefront-per-circuit kernel once per element-kind set, each in its own
# translation unit.
PERSIST_WAVES = (1, 2, 4)
PERSIST_KINDS = (126, 190, 318, 32766)  # diodes, level-1 MOSFETs, BJTs, everything (kh_persist.h)
UNITS = [
    ("kronhip.hip", [], "kronhip.o"),
    ("kh_symbolic.cpp", [], "kh_symbolic.o"),
    ("kh_output.cpp", [], "kh_output.o"),
    ("kh_mlp.hip", [], "kh_mlp.o"),
] + [("kh_persist.hip", [f"-DKH_PERSIST_WAVES={w}", f"-DKH_KINDS={k}"], f"kh_persist_w{w}_k{k}.o")
     for w in PERSIST_WAVES for k in PERSIST_KINDS] + [
    ("kh_wave.hip", [f"-DKH_KINDS={k}"], f"kh_wave_k{k}.o") for k in PERSIST_KINDS]
BENCH = os.path.join(HERE, "kh_bench")
OBJ = os.path.join(HERE, "_obj")


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc ({what}) failed:\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src, extra, obj in UNITS:
        cmd = [HIPCC] + FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(OBJ, obj)]
        jobs.append((cmd, src + " " + " ".join(extra)))
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 8)) as ex:
        results = list(ex.map(lambda j: _run(*j), jobs))
    if verbose:
        for r in results:
            if r.stderr:
                print(r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950"] + [os.path.join(OBJ, o) for _, _, o in UNITS] +
         ["-o", tmp], "link")
    os.replace(tmp, LIB)
    # kh_bench: the C++ driver (and one copy kernel), linked against the library next to it.
    tmpb = BENCH + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-x", "hip", os.path.join(HERE, "kh_bench.cpp"), "-x", "none", "-L" + HERE, "-lkronhip",
          "-Wl,-rpath,$ORIGIN", "-o", tmpb], "kh_bench")
    os.replace(tmpb, BENCH)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
