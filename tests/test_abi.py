"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/kronhip.h declares (no compute without a GPU), and
the host-side setup (netlist, builder, batch generators) behaves as the
reference's."""
import ctypes
import os
import re

import numpy as np
import pytest

from kronspice_amd import api, circuit as kc, netlist, workloads as W
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "kronhip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(kh_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libkronhip.so not built")
    lib = ctypes.CDLL(api.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_node_numbering_follows_builder():
    # circuit/builder.go:38-48: order of first appearance; MOSFET terminals
    # are registered drain, source, gate, bulk (circuit/builder.go:113-118).
    d = netlist.parse("t\nV1 a 0 1\nM1 d g s b nm\nR1 a d 1k\n.model nm NMOS(LEVEL=1)\n.op\n.end\n")
    assert d.circuit.node_names == ["0", "a", "d", "s", "g", "b"]


def test_parse_value_suffixes():
    # netlist/netlist_test.go:18 and netlist/units.go:17-23
    for s, v in [("1meg", 1e6), ("1MEG", 1e6), ("1m", 1e-3), ("1M", 1e-3), ("2u", 2e-6), ("10u", 1e-5),
                 ("4.7kohm", 4700.0), ("159.15494n", 159.15494e-9), ("1mil", 25.4e-6)]:
        assert netlist.parse_value(s) == v, s


def test_ladder_batch_shape():
    b = W.resistor_ladders(B=8)
    assert b.n == 65  # 64 nodes + the source branch
    els, p = b.elements, b.circuit_params(3)
    st, x, info = oracle.op(b.circuit.num_nodes, els, p)
    assert st == 0 and info["nnz_a"] == 192  # 64 diagonal + 126 series + 2 branch
    # Regeneration is per circuit (counter-based): circuit 3 alone equals circuit 3 of a larger batch.
    b2 = W.resistor_ladders(B=64)
    assert np.array_equal(b2.circuit_params(3), p)


def test_rectifier_batch_runs_in_oracle():
    b = W.rectifier_rc_ladders(B=2)
    assert b.n == 35
    spec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], start=0.0, max_step=b.spec["max_step"])
    x, st, acc, rej, led = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params, b.n, spec,
                                             lane_spec=b.lane_spec)
    assert (st == 0).all() and (acc >= 2000).all()


def test_ctypes_structures_match_the_header(tmp_path):
    """The Python mirror's structures (api.Options, Stats, TranSpec, Group,
    DcSpec; the oracle's KhOptions, KhTranSpec) have the C header's sizes and
    field offsets, so neither side reads past or short of the other's."""
    import subprocess
    src = tmp_path / "sz.c"
    fields = {
        "kh_options": [f for f, _ in api.Options._fields_],
        "kh_stats": [f for f, _ in api.Stats._fields_],
        "kh_tran_spec": [f for f, _ in api.TranSpec._fields_],
        "kh_group": [f for f, _ in api.Group._fields_],
        "kh_dc_spec": [f for f, _ in api.DcSpec._fields_],
        "kh_candidate": [f for f, _ in api.Candidate._fields_],
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{ROOT}/include/kronhip.h"', "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'  printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            base = f.split("[")[0]
            lines.append(f'  printf("{st}.{base} %zu\\n", offsetof({st}, {base}));')
    lines.append("  return 0;\n}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "sz"
    r = subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    py = {"kh_options": api.Options, "kh_stats": api.Stats, "kh_tran_spec": api.TranSpec, "kh_group": api.Group,
          "kh_dc_spec": api.DcSpec, "kh_candidate": api.Candidate}
    for st, cls in py.items():
        assert int(got[st]) == ctypes.sizeof(cls), st
        for f, _ in cls._fields_:
            assert int(got[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)
    assert ctypes.sizeof(oracle.KhOptions) == int(got["kh_options"])
    assert ctypes.sizeof(oracle.KhTranSpec) == int(got["kh_tran_spec"])
