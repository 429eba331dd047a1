"""The reference's result formats written by the library (kh_format_op,
kronspice_amd/csrc/kh_output.cpp): the `kronspice -json` document
(output/json.go:18-316 through encoding/json with two-space indentation) and
output.WriteCSV's rows (output/csv.go:15-48 through encoding/csv).

CPU (the library's host side, no device):
  * whole documents byte for byte against texts written out by hand from the
    formats' rules: field order and names, omitted empty fields, sorted node
    keys, two-space indentation, the trailing newline; the ledger rows of
    every port a device declares, with their notes; the supplied list only
    when an element supplies power;
  * encoding/json's number and string rules at their edges: 'f' notation
    from 1e-6 up to 1e21, 'e' outside with the exponent's leading zero
    dropped, shortest round-trip digits, -0; HTML-safe escapes, control
    characters, U+2028/U+2029, invalid UTF-8, non-ASCII kept raw;
  * encoding/csv quoting and strconv 'g' 12 digits; the reference's own CSV
    expectations (output/output_test.go:24-36);
  * the failed run's document (cmd/kronspice/main.go:127-140, :284-305).

GPU: op(report=True) on golden decks, written through kh_format_op, parses
back to the oracle's report within 1e-9 (bit-identical bytes where the
values are).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from kronspice_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def hostlib():
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libkronhip.so not built")
    return ctypes.CDLL(api.LIB_PATH)


def elems(spec):
    """[(kind, n0, n1)] -> element table."""
    from kronspice_amd.circuit import ELEMENT_DTYPE
    e = np.zeros(len(spec), dtype=ELEMENT_DTYPE)
    for k, (kind, a, b) in enumerate(spec):
        e[k]["kind"] = kind
        e[k]["node"] = [a, b, 0, 0]
    return e


R, C_, L_, V_, I_, D, M, Q = 1, 2, 3, 4, 5, 6, 7, 8

DIVIDER_NODES = ["0", "in", "out"]
DIVIDER_NAMES = ["V1", "R1", "R2"]
DIVIDER_ELEMS = elems([(V_, 1, 0), (R, 1, 2), (R, 2, 0)])
DIVIDER_X = np.array([10.0, 5.0, -0.005])
DIVIDER_REP = np.array([[10, -0.005, -0.05, 0, 0, 0, 0.05, 0],
                        [5, 0.005, 0.025, 0, 0, 0, -0.025, 2],
                        [5, 0.005, 0.025, 0, 0, 0, -0.025, 2]], dtype=np.float64)
DIVIDER_LED = np.array([0.05, 0, 0.05, 0, 0, 0, 0, 0.1, 1e-17, 0])


def element_json(name, kind, a, b, v, i, p, w=0, ms=0, ind="          "):
    return (f'{ind}{{\n'
            f'{ind}  "name": "{name}",\n'
            f'{ind}  "kind": "{kind}",\n'
            f'{ind}  "nodes": [\n{ind}    "{a}",\n{ind}    "{b}"\n{ind}  ],\n'
            f'{ind}  "voltage": {v},\n'
            f'{ind}  "current": {i},\n'
            f'{ind}  "power_absorbed": {p},\n'
            f'{ind}  "stored_energy": {w},\n'
            f'{ind}  "model_supplied": {ms}\n'
            f'{ind}}}')


def row_json(el, ch, pw, note, ind="            "):
    return (f'{ind}{{\n'
            f'{ind}  "element": "{el}",\n'
            f'{ind}  "channel": "{ch}",\n'
            f'{ind}  "status": "modeled",\n'
            f'{ind}  "power_w": {pw},\n'
            f'{ind}  "note": "{note}"\n'
            f'{ind}}}')


INSIDE = "terminals connect inside the network"
JOULE = "Joule heating i²R to surroundings"
IMPRESSED = "impressed source: exchanges energy with the exterior"

DIVIDER_JSON = (
    '{\n'
    '  "tool": "kronspice",\n'
    '  "version": "dev",\n'
    '  "ok": true,\n'
    '  "title": "Divider",\n'
    '  "analyses": [\n'
    '    {\n'
    '      "kind": "op",\n'
    '      "op": {\n'
    '        "nodes": {\n'
    '          "in": 10,\n'
    '          "out": 5\n'
    '        },\n'
    '        "elements": [\n'
    + element_json("V1", "voltage source", "in", "0", 10, -0.005, -0.05) + ',\n'
    + element_json("R1", "resistor", "in", "out", 5, 0.005, 0.025) + ',\n'
    + element_json("R2", "resistor", "out", "0", 5, 0.005, 0.025) + '\n'
    '        ],\n'
    '        "ledger": {\n'
    '          "inflow": 0.05,\n'
    '          "outflow": 0,\n'
    '          "dissipation": 0.05,\n'
    '          "storage_rate": 0,\n'
    '          "model_supplied": 0,\n'
    '          "hypothetical": 0,\n'
    '          "residual": 0,\n'
    '          "relative_residual": 0,\n'
    '          "rows": [\n'
    + row_json("V1", "electrical", 0.05, IMPRESSED) + ',\n'
    + row_json("R1", "electrical", 0, INSIDE) + ',\n'
    + row_json("R1", "thermal", -0.025, JOULE) + ',\n'
    + row_json("R2", "electrical", 0, INSIDE) + ',\n'
    + row_json("R2", "thermal", -0.025, JOULE) + '\n'
    '          ]\n'
    '        }\n'
    '      }\n'
    '    }\n'
    '  ]\n'
    '}\n')

DIVIDER_CSV = (
    'name,value\n'
    'v(in),10\n'
    'v(out),5\n'
    'i(v1),-0.005\n'
    'p(v1),-0.05\n'
    'i(r1),0.005\n'
    'p(r1),0.025\n'
    'i(r2),0.005\n'
    'p(r2),0.025\n'
    '"ledger(v1,electrical)",0.05\n'
    '"ledger(r1,electrical)",0\n'
    '"ledger(r1,thermal)",-0.025\n'
    '"ledger(r2,electrical)",0\n'
    '"ledger(r2,thermal)",-0.025\n'
    'ledger.inflow,0.05\n'
    'ledger.outflow,0\n'
    'ledger.dissipation,0.05\n'
    'ledger.residual,0\n')


def fmt(fmt_, status=0, x=DIVIDER_X, rep=DIVIDER_REP, led=DIVIDER_LED, nodes=DIVIDER_NODES, names=DIVIDER_NAMES,
        els=DIVIDER_ELEMS, title="Divider", version=None):
    return api.format_op(nodes, names, els, status, x, rep, led, fmt=fmt_, title=title, version=version,
                         library=hostlib()).decode()


def test_divider_document_bytes():
    assert fmt(api.FORMAT_JSON) == DIVIDER_JSON
    doc = json.loads(DIVIDER_JSON)  # the expected text is a valid document of the reference's shape
    assert doc["analyses"][0]["op"]["nodes"]["out"] == 5


def test_divider_csv_bytes():
    assert fmt(api.FORMAT_CSV) == DIVIDER_CSV


def test_reference_csv_expectations():
    """output/output_test.go:24-36: a node a at 2 V, R1 from a to ground with
    1 A and 2 W -> "v(a),2" and "p(r1),2" rows."""
    out = fmt(api.FORMAT_CSV, x=np.array([2.0]), rep=np.array([[2, 1, 2, 0, 0, 0, -2, 2]], np.float64),
              led=np.zeros(10), nodes=["0", "a"], names=["R1"], els=elems([(R, 1, 0)]), title="t")
    assert "v(a),2\n" in out and "p(r1),2\n" in out


@pytest.mark.parametrize("v,text", [
    (0.0, "0"), (-0.0, "-0"), (1.0, "1"), (-2.5, "-2.5"), (1e-6, "0.000001"), (9.99e-7, "9.99e-7"),
    (1e-7, "1e-7"), (2.5e-10, "2.5e-10"), (1.5e-100, "1.5e-100"), (5e-324, "5e-324"),
    (1e20, "100000000000000000000"), (1e21, "1e+21"), (-1.2345e22, "-1.2345e+22"),
    (0.1 + 0.2, "0.30000000000000004"), (123456.789, "123456.789"), (1 / 3, "0.3333333333333333"),
    (1.7976931348623157e308, "1.7976931348623157e+308"), (4.999999999999693e-17, "4.999999999999693e-17"),
])
def test_json_numbers(v, text):
    out = fmt(api.FORMAT_JSON, x=np.array([v, 0.0, 0.0]))
    assert f'          "in": {text},\n' in out, out[:400]


@pytest.mark.parametrize("title,text", [
    ("a<b>&c", r"a\u003cb\u003e\u0026c"),
    ('q"b\\s', r'q\"b\\s'),
    ("t\tn\nr\rb\bf\f", r"t\tn\nr\rb\bf\f"),
    ("\x01\x1f", r"\u0001\u001f"),
    ("½µ→€𝄞", "½µ→€𝄞"),
    ("l\u2028p\u2029", r"l\u2028p\u2029"),
])
def test_json_strings(title, text):
    out = fmt(api.FORMAT_JSON, title=title)
    assert f'  "title": "{text}",\n' in out


def test_json_invalid_utf8_replaced():
    L = hostlib()
    L.kh_format_op.restype = ctypes.c_int64
    nn = (ctypes.c_char_p * 3)(b"0", b"in", b"out")
    en = (ctypes.c_char_p * 3)(b"V1", b"R1", b"R2")
    nm = api.Names(title=b"x\xffy\xc3", num_nodes=3, n_elems=3, node_names=nn, elem_names=en,
                   elems=DIVIDER_ELEMS.ctypes.data)
    ptr = [np.ascontiguousarray(a).ctypes.data for a in (DIVIDER_X, DIVIDER_REP, DIVIDER_LED)]
    buf = ctypes.create_string_buffer(8192)
    n = L.kh_format_op(0, ctypes.byref(nm), None, 0, *[ctypes.c_void_p(p) for p in ptr], buf, 8192)
    assert b'"title": "x\\ufffdy\\ufffd",' in buf.raw[:n]


def test_empty_title_omitted_and_version():
    out = fmt(api.FORMAT_JSON, title="", version="v1.2.3")
    assert '"title"' not in out and out.startswith('{\n  "tool": "kronspice",\n  "version": "v1.2.3",\n  "ok": true,\n'
                                                   '  "analyses": [\n')


def test_node_keys_sorted_bytewise():
    nodes = ["0", "b", "B", "a10", "a9", "_x"]
    x = np.arange(1, 6, dtype=np.float64)
    out = fmt(api.FORMAT_JSON, x=x, nodes=nodes, names=[], els=elems([]), rep=np.zeros((0, 8)))
    block = out[out.index('"nodes": {'):]
    block = block[:block.index("}")]
    keys = [ln.split('"')[1] for ln in block.splitlines()[1:] if '"' in ln]
    assert keys == ["B", "_x", "a10", "a9", "b"]
    assert '"elements": null' in out and '"rows": []' in out


def test_negative_resistance_supplied_list():
    nodes, names = ["0", "a", "b"], ["V1", "R1", "Rn"]
    els = elems([(V_, 1, 0), (R, 1, 2), (R, 2, 0)])
    rep = np.array([[1, -0.004, -0.004, 0, 0, 0, 0.004, 0],
                    [0.2, 0.004, 0.0008, 0, 0, 0, -0.0008, 2],
                    [0.8, -0.004, -0.0032, 0, 0, 0.0032, 0, -1]], np.float64)
    led = np.array([0.004, 0, 0.0008, 0, 0.0032, 0, 0.0064, 0.0112, 0, 0.5714285714285714])
    out = fmt(api.FORMAT_JSON, x=np.array([1.0, 0.8, -0.004]), rep=rep, led=led, nodes=nodes, names=names, els=els,
              title="neg")
    assert row_json("Rn", "electrical", 0, "idealized negative resistance: delivers v²/|R| as model-supplied power") \
        in out
    assert '"Rn",\n              "channel": "thermal"' not in out
    assert out.endswith('          "supplied": [\n'
                        '            {\n'
                        '              "Element": "Rn",\n'
                        '              "PowerW": 0.0032\n'
                        '            }\n'
                        '          ]\n'
                        '        }\n      }\n    }\n  ]\n}\n')


def test_junction_device_rows_and_kinds():
    nodes, names = ["0", "a", "b", "c"], ["D1", "M1", "Q1", "C1", "L1", "I1"]
    els = elems([(D, 1, 0), (M, 1, 2), (Q, 3, 2), (C_, 1, 0), (L_, 2, 0), (I_, 3, 0)])
    rep = np.zeros((6, 8))
    rep[:, 7] = [2, 2, 2, -1, -1, 0]
    out = fmt(api.FORMAT_JSON, x=np.zeros(4), rep=rep, led=np.zeros(10), nodes=nodes, names=names, els=els)
    for k in ("diode", "MOSFET", "BJT", "capacitor", "inductor", "current source"):
        assert f'"kind": "{k}"' in out
    for n in ("terminals connect inside the network; stores junction charge",
              "junction and series-resistance heating to surroundings",
              "terminals connect inside the network; stores gate and junction charge",
              "channel, junction and series-resistance heating to surroundings",
              "junction, transport and series-resistance heating to surroundings",
              "terminals connect inside the network; stores ½Cv²",
              "terminals connect inside the network; stores ½Li²"):
        assert f'"note": "{n}"' in out
    assert out.count('"status": "modeled"') == 2 + 2 + 2 + 1 + 1 + 1


@pytest.mark.parametrize("status,cls", [(api.NOCONV, "no-convergence"), (api.SINGULAR, "singular"),
                                        (api.LEDGER, "ledger"), (api.TSTEP_TOO_SMALL, "no-convergence")])
def test_failed_run_document(status, cls):
    out = fmt(api.FORMAT_JSON, status=status, x=None, rep=None, led=None)
    doc = json.loads(out)
    assert doc["ok"] is False and doc["error"]["class"] == cls and doc["analyses"] == []
    assert out.startswith('{\n  "tool": "kronspice",\n  "version": "dev",\n  "ok": false,\n  "error": {\n'
                          f'    "class": "{cls}",\n')
    assert out.endswith('  "title": "Divider",\n  "analyses": []\n}\n')
    assert fmt(api.FORMAT_CSV, status=status, x=None, rep=None, led=None) == ""


@pytest.mark.parametrize("v,text", [(1 / 3, "0.333333333333"), (1e-13, "1e-13"), (1234567890123456.0, "1.23456789012e+15"),
                                    (123456789012.0, "123456789012"), (1e-4, "0.0001"), (1e-5, "1e-05"),
                                    (float("nan"), "NaN"), (float("inf"), "+Inf"), (-float("inf"), "-Inf")])
def test_csv_numbers(v, text):
    out = fmt(api.FORMAT_CSV, x=np.array([v, 5.0, 0.0]))
    assert f"v(in),{text}\n" in out


def test_csv_quotes_fields():
    out = fmt(api.FORMAT_CSV, nodes=["0", 'a"b', " sp"], title="q")
    assert '"v(a""b)",10\n' in out and 'v( sp),5\n' in out


def test_bad_arguments():
    L = hostlib()
    L.kh_format_op.restype = ctypes.c_int64
    assert L.kh_format_op(0, None, None, 0, None, None, None, None, 0) == -1
    with pytest.raises(api.KhError):
        fmt(api.FORMAT_JSON, x=None)


# ---- GPU: the device's report written out ----------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["divider", "bridge", "ladder", "diode_fwd", "mos1_op", "bjt_ce_op", "two_sources",
                                  "isource", "zener_dc"])
def test_gpu_report_documents_match_oracle(name):
    from oracle import oracle
    from tests import golden_runner as gr
    deck = gr.load_deck(name)
    c = deck.circuit
    els, p = c.flatten()
    t = api.Topology(c.num_nodes, els, p)
    B = 4
    r = api.op(t, np.repeat(p[:, None], B, axis=1), report=True)
    st = r.status.cpu().numpy()
    x = r.x.cpu().numpy()
    rep = r.report.cpu().numpy()
    led = r.report_ledger.cpu().numpy()
    e0, x0, rep0, led0 = oracle.op_report(c.num_nodes, els, p)
    assert (st == e0).all() and e0 == 0
    names = [e.name for e in c.elements]
    want = api.format_op(c.node_names, names, els, int(e0), x0, rep0, led0, title=c.title)
    for b in range(B):
        got = api.format_op(c.node_names, names, els, int(st[b]), x[:, b], rep[b], led[b], title=c.title)
        if np.array_equal(rep[b], rep0) and np.array_equal(led[b], led0) and np.array_equal(x[:, b], x0):
            assert got == want
        dg, dw = json.loads(got), json.loads(want)
        og, ow = dg["analyses"][0]["op"], dw["analyses"][0]["op"]
        assert og["nodes"].keys() == ow["nodes"].keys()
        for eg, ew in zip(og["elements"], ow["elements"]):
            for f in ("voltage", "current", "power_absorbed", "stored_energy", "model_supplied"):
                assert abs(eg[f] - ew[f]) <= 1e-9 * max(abs(ew[f]), 1e-12), (eg["name"], f, eg[f], ew[f])
        for f in ("inflow", "outflow", "dissipation", "storage_rate", "model_supplied"):
            assert abs(og["ledger"][f] - ow["ledger"][f]) <= 1e-9 * max(abs(ow["ledger"][f]), 1e-15)
        assert len(og["ledger"]["rows"]) == len(ow["ledger"]["rows"])
        csv_g = api.format_op(c.node_names, names, els, int(st[b]), x[:, b], rep[b], led[b], fmt=api.FORMAT_CSV)
        csv_w = api.format_op(c.node_names, names, els, int(e0), x0, rep0, led0, fmt=api.FORMAT_CSV)
        assert [ln.split(",")[0] for ln in csv_g.decode().splitlines()] == \
               [ln.split(",")[0] for ln in csv_w.decode().splitlines()]


@pytest.mark.gpu
def test_gpu_report_batch_against_oracle():
    """The device's element reports and ledgers over a batch of 256 two-stage
    BJT amplifiers (polished operating points) against the oracle's at 1e-9."""
    from kronspice_amd import workloads as W
    from oracle import oracle
    b, _ = W.bjt_amplifiers(256)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.op(t, b.params, polish=True, report=True)
    st = r.status.cpu().numpy()
    rep = r.report.cpu().numpy()
    led = r.report_ledger.cpu().numpy()
    for c in range(0, 256, 16):
        e0, x0, rep0, led0 = oracle.op_report(b.circuit.num_nodes, b.elements, np.ascontiguousarray(b.params[:, c]),
                                               polish=True)
        assert st[c] == e0 == 0
        scale = np.maximum(np.abs(rep0), 1e-18)
        assert (np.abs(rep[c] - rep0) <= 1e-9 * scale + 1e-15 * np.abs(rep0).max(axis=0)).all(), c
        assert np.array_equal(rep[c][:, 7], rep0[:, 7])
        for f in range(6):
            assert abs(led[c][f] - led0[f]) <= 1e-9 * max(abs(led0[f]), 1e-15), (c, f)


# ---- DC sweeps --------------------------------------------------------------------------------
SWEEP_NODES = ["0", "in", "out"]
SWEEP_NAMES = ["V1", "R1", "R2"]
SWEEP_ELEMS = elems([(V_, 1, 0), (R, 1, 2), (R, 2, 0)])


def sweep_arrays(values):
    """The reference's dcSweep deck (cmd/kronspice/main_test.go:44-50):
    in = V, out = V/2, every current V/2000 in magnitude."""
    x = np.array([[v, v / 2, -v / 2000] for v in values])
    i = np.array([[-v / 2000, v / 2000, v / 2000] for v in values])
    return x, i


def series(name, vals, ind="          "):
    body = ",\n".join(f"{ind}  {v}" for v in vals)
    return f'{ind}"{name}": [\n{body}\n{ind}]'


def dc_fmt(fmt_, status=0, sweeps=((0, 0.0, 2.0, 1.0),), sources=("V1",), values=(0.0, 1.0, 2.0), max_rel=0.0):
    x, i = sweep_arrays(values)
    return api.format_dc(SWEEP_NODES, SWEEP_NAMES, SWEEP_ELEMS, status, 3, list(sweeps), list(sources), len(values),
                         x, i, max_rel, fmt=fmt_, title="Sweep", library=hostlib()).decode()


def test_dc_document_bytes():
    want = ('{\n  "tool": "kronspice",\n  "version": "dev",\n  "ok": true,\n  "title": "Sweep",\n'
            '  "analyses": [\n    {\n      "kind": "dc",\n      "dc": {\n'
            '        "sources": [\n          "V1"\n        ],\n'
            '        "values": [\n          [\n            0,\n            1,\n            2\n          ]\n        ],\n'
            '        "v": {\n' + series("in", [0, 1, 2]) + ',\n' + series("out", [0, 0.5, 1]) + '\n        },\n'
            '        "i": {\n' + series("r1", [0, 0.0005, 0.001]) + ',\n' + series("r2", [0, 0.0005, 0.001]) + ',\n'
            + series("v1", ["-0", -0.0005, -0.001]) + '\n        },\n'
            '        "max_relative_residual": 0\n      }\n    }\n  ]\n}\n')
    got = dc_fmt(api.FORMAT_JSON)
    assert got == want
    doc = json.loads(got)["analyses"][0]["dc"]
    assert doc["v"]["out"] == [0, 0.5, 1]  # the reference's expectation (main_test.go:151-157)


def test_dc_csv_bytes():
    assert dc_fmt(api.FORMAT_CSV) == ("v1,v(in),v(out),i(v1),i(r1),i(r2)\n"
                                      "0,0,0,-0,0,0\n"
                                      "1,1,0.5,-0.0005,0.0005,0.0005\n"
                                      "2,2,1,-0.001,0.001,0.001\n")


def test_dc_two_sweeps_values():
    """Two sources: the inner sweep's values repeat for every outer value,
    values[1] holds the outer value of each point (analysis/dc.go:104-119)."""
    vals = [0.0, 0.5, 1.0] * 2
    out = dc_fmt(api.FORMAT_JSON, sweeps=((0, 0.0, 1.0, 0.5), (0, 3.0, 4.0, 1.0)), sources=("V1", "V2"),
                 values=vals)
    doc = json.loads(out)["analyses"][0]["dc"]
    assert doc["sources"] == ["V1", "V2"]
    assert doc["values"] == [[0, 0.5, 1, 0, 0.5, 1], [3, 3, 3, 4, 4, 4]]
    csv = dc_fmt(api.FORMAT_CSV, sweeps=((0, 0.0, 1.0, 0.5), (0, 3.0, 4.0, 1.0)), sources=("V1", "V2"), values=vals)
    assert csv.splitlines()[0] == "v1,v2,v(in),v(out),i(v1),i(r1),i(r2)" and csv.splitlines()[4].startswith("0,4,")


def test_dc_failed_sweep_document():
    out = dc_fmt(api.FORMAT_JSON, status=api.NOCONV)
    doc = json.loads(out)
    assert doc["ok"] is False and doc["error"]["class"] == "no-convergence" and doc["analyses"] == []


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["diode_clipper_dc", "bjt_output_dc", "mos1_idvd_dc", "zener_dc"])
def test_gpu_dc_currents_and_document(name):
    """dc(currents=True) on golden sweeps: each point's element currents
    against the oracle's sweep (the same warm-started points) at 1e-9, and
    the document's series equal to the arrays it was written from."""
    from oracle import oracle
    from tests import golden_runner as gr
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "dc"][0]
    els, p = c.flatten()
    sw = [(c.element_index(a.args["source"]), a.args["start"], a.args["stop"], a.args["step"])]
    if a.args.get("source2"):
        sw.append((c.element_index(a.args["source2"]), a.args["start2"], a.args["stop2"], a.args["step2"]))
    t = api.Topology(c.num_nodes, els, p)
    r = api.dc(t, p[:, None], sw, currents=True)
    st, pts = int(r.status.cpu()[0]), int(r.points.cpu()[0])
    e0, x0, it0, d0, mr0, i0 = oracle.dc(c.num_nodes, els, p, sw, currents=True)
    assert st == e0 == 0 and pts == d0
    cur = r.currents.cpu().numpy()[:, :, 0]
    x = r.x.cpu().numpy()[:, :, 0]
    scale = np.maximum(np.abs(i0).max(axis=0), 1e-15)
    assert (np.abs(cur[:pts] - i0[:pts]) <= 1e-9 * scale).all()
    names = [e.name for e in c.elements]
    srcs = [a.args["source_name"]]
    doc = json.loads(api.format_dc(c.node_names, names, els, st, t.n, sw, srcs, pts, x[:pts], cur[:pts],
                                   float(r.ledger.cpu()[0]), title=c.title))
    dcd = doc["analyses"][0]["dc"]
    for k, e in enumerate(c.elements):
        assert dcd["i"][e.name.lower()] == cur[:pts, k].tolist()
    for i, nn in enumerate(c.node_names[1:]):
        assert dcd["v"][nn] == x[:pts, i].tolist()


# ---- AC sweeps --------------------------------------------------------------------------------
AC_NODES = ["0", "out"]
AC_NAMES = ["V1", "R1"]
AC_ELEMS = elems([(V_, 1, 0), (R, 1, 0)])
AC_FREQS = np.array([10.0, 100.0, 1000.0, 1e4])
# phasors with exact magnitudes and phases: 0°, 90°, 180°, −90°
AC_X = np.array([[1 + 0j, -0.001 + 0j], [2j, -0.002j], [-0.5 + 0j, 0.0005 + 0j], [-4j, 0.004j]])  # out, V1 branch
AC_I = np.array([[-0.001 + 0j, 0.001 + 0j], [-0.002j, 0.002j], [0.0005 + 0j, -0.0005 + 0j], [0.004j, -0.004j]])
AC_LED = np.array([[0.0005, 0, 0.0005, 0, 0, 0], [0.002, 0, 0.002, 0, 0, 0],
                   [0.000125, 0, 0.000125, 0, 0, 0], [0.008, 0, 0.008, 0, 0, 0]])


def ac_fmt(fmt_, status=0):
    return api.format_ac(AC_NODES, AC_NAMES, AC_ELEMS, status, 2, AC_FREQS, AC_X, AC_I, AC_LED, 0.0, fmt=fmt_,
                         title="ac", library=hostlib()).decode()


def phasor_json(name, mags, phases, ind="          "):
    m = ",\n".join(f"{ind}    {v}" for v in mags)
    p = ",\n".join(f"{ind}    {v}" for v in phases)
    return (f'{ind}"{name}": {{\n{ind}  "mag": [\n{m}\n{ind}  ],\n'
            f'{ind}  "phase_deg": [\n{p}\n{ind}  ]\n{ind}}}')


def test_ac_document_bytes():
    ind = "          "
    led = lambda name, vals: f'{ind}"{name}": [\n' + ",\n".join(f"{ind}  {v}" for v in vals) + f"\n{ind}]"
    want = ('{\n  "tool": "kronspice",\n  "version": "dev",\n  "ok": true,\n  "title": "ac",\n'
            '  "analyses": [\n    {\n      "kind": "ac",\n      "ac": {\n'
            '        "freq": [\n          10,\n          100,\n          1000,\n          10000\n        ],\n'
            '        "v": {\n' + phasor_json("out", [1, 2, 0.5, 4], [0, 90, 180, -90]) + '\n        },\n'
            '        "i": {\n' + phasor_json("r1", [0.001, 0.002, 0.0005, 0.004], [0, 90, 180, -90]) + ',\n'
            + phasor_json("v1", [0.001, 0.002, 0.0005, 0.004], [180, -90, 0, 90]) + '\n        },\n'
            '        "ledger": {\n'
            + led("inflow", [0.0005, 0.002, 0.000125, 0.008]) + ',\n'
            + led("outflow", [0, 0, 0, 0]) + ',\n'
            + led("dissipation", [0.0005, 0.002, 0.000125, 0.008]) + ',\n'
            + led("reactive", [0, 0, 0, 0]) + ',\n'
            + led("model_supplied", [0, 0, 0, 0]) + ',\n'
            + led("residual", [0, 0, 0, 0]) + ',\n'
            '          "max_relative_residual": 0\n        }\n      }\n    }\n  ]\n}\n')
    got = ac_fmt(api.FORMAT_JSON)
    assert got == want


def test_ac_csv_bytes():
    assert ac_fmt(api.FORMAT_CSV) == ("freq,vm(out),vp(out),ledger.inflow,ledger.dissipation,ledger.reactive\n"
                                      "10,1,0,0.0005,0.0005,0\n"
                                      "100,2,90,0.002,0.002,0\n"
                                      "1000,0.5,180,0.000125,0.000125,0\n"
                                      "10000,4,-90,0.008,0.008,0\n")


def test_ac_reference_expectation_shape():
    """cmd/kronspice/main_test.go:125-145: nine frequencies of .ac dec 2 10 100k,
    the corner at index 4 (1 kHz), a dissipation point per frequency."""
    from kronspice_amd import netlist
    freqs = netlist.ac_frequencies("dec", 2, 10, 1e5)
    assert len(freqs) == 9 and abs(freqs[4] - 1000) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rc_lowpass_ac", "rlc_bandpass_ac", "bjt_ce_ac", "mos1_cs_ac", "mos3_cs_ac"])
def test_gpu_ac_currents_ledgers_and_document(name):
    """ac(currents=True) against the oracle's AC sweep: element currents and
    per-frequency ledgers at 1e-9 (the solutions bit-identical where the
    operating point involves no transcendental of the device library, as the
    AC parity tests have it: not for the BJT and level-3 MOSFET decks); the
    document written from them parses back to the same arrays."""
    from kronspice_amd import netlist
    from oracle import oracle
    from tests import golden_runner as gr
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "ac"][0]
    freqs = np.asarray(netlist.ac_frequencies(a.args["variation"], a.args["points"], a.args["fstart"],
                                              a.args["fstop"]), np.float64)
    els, p = c.flatten()
    t = api.Topology(c.num_nodes, els, p)
    r = api.ac(t, p[:, None], freqs, currents=True)
    st = int(r.status.cpu()[0])
    e0, X0, mr0, cur0, led0 = oracle.ac(c.num_nodes, els, p, freqs, currents=True)
    assert st == e0 == 0
    X = r.x.cpu().numpy()[:, :, 0]
    cur = r.currents.cpu().numpy()[:, :, 0]
    led = r.ledgers.cpu().numpy()[:, :, 0]
    if name in ("rc_lowpass_ac", "rlc_bandpass_ac", "mos1_cs_ac"):
        assert np.array_equal(X, X0)
    else:
        assert (np.abs(X - X0) <= 1e-9 * np.maximum(np.abs(X0).max(axis=0), 1e-300)).all()
    sc = np.maximum(np.abs(cur0).max(axis=0), 1e-18)
    assert (np.abs(cur - cur0) <= 1e-9 * sc).all()
    # every ledger term against the sweep's largest power (the residual and
    # tiny reactive terms are round-off themselves)
    assert (np.abs(led - led0) <= 1e-9 * max(np.abs(led0).max(), 1e-300)).all()
    names = [e.name for e in c.elements]
    doc = json.loads(api.format_ac(c.node_names, names, els, st, t.n, freqs, X, cur, np.ascontiguousarray(led),
                                   float(r.ledger.cpu()[0]), title=c.title))
    acd = doc["analyses"][0]["ac"]
    assert acd["freq"] == freqs.tolist()
    for k, e in enumerate(c.elements):
        got = np.array(acd["i"][e.name.lower()]["mag"])
        assert np.allclose(got, np.abs(cur[:, k]), rtol=1e-15, atol=0)
    assert acd["ledger"]["dissipation"] == led[:, 2].tolist()
