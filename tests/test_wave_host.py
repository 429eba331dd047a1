"""The wavefront-per-circuit kernel's device code (kronspice_amd/csrc/kh_wave.h)
on the CPU: compiled with g++ against a host emulation of the HIP built-ins
it uses (tests/host_shim/hip/hip_runtime.h: the 64 lanes of a wavefront are
64 fibers meeting at every barrier, ballot and shuffle), run by
tests/wave_host.cpp beside the lane-per-circuit kernel's code (run_circuit,
kh_group.h) on the same circuits, with the host's pattern analysis between
launches. Bar: every circuit's status, Newton iterations, accepted and
rejected steps, solution and ledger bit-identical between the two."""
import os
import subprocess

import pytest

from kronspice_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("wave") / "wave_host")
    csrc = os.path.join(ROOT, "kronspice_amd", "csrc")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-DKH_HD=", "-I", os.path.join(ROOT, "tests", "host_shim"),
           "-I", csrc, os.path.join(ROOT, "tests", "wave_host.cpp"), os.path.join(csrc, "kh_symbolic.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def run(exe, tmp_path, batches, transient, per, backend=0, max_solves=7):
    path = str(tmp_path / "batch.bin")
    W.write_batch_file(path, batches, transient=transient)
    r = subprocess.run([exe, path, str(per), str(backend), str(max_solves)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    return r.stdout


def test_wave_ladders_op(exe, tmp_path):
    run(exe, tmp_path, [W.resistor_ladders(8)], False, 4, backend=2)


def test_wave_rectifier_transient(exe, tmp_path):
    run(exe, tmp_path, [W.rectifier_rc_ladders(4, steps=100)], True, 4)


def test_wave_rectifier_transient_dense_backend(exe, tmp_path):
    run(exe, tmp_path, [W.rectifier_rc_ladders(2, steps=40)], True, 2, backend=1)


def test_wave_bjt_amplifier_op(exe, tmp_path):
    b, _ = W.bjt_amplifiers(4)
    run(exe, tmp_path, [b], False, 4)


def test_wave_mosfet_chains_uneven_transient(exe, tmp_path):
    bs = [W.mosfet_chain_batch(p, 1, 3, steps=12) for p in (33, 91, 151, 199, 17, 120)]
    run(exe, tmp_path, bs, True, 1, backend=2)
