"""Linear controlled sources E (VCVS), F (CCCS), G (VCCS), H (CCVS)
(device/controlled.go:13-161).

Oracle pins (CPU): the reference's controlled_op golden (exact nodal values,
run with the other goldens in tests/test_oracle.py), the node order of
circuit.FromDeck (control nodes registered before output nodes), and closed
forms — each source driving an RC load in AC (the gain times the load's
transfer function, the H and F loops through their sensing source), an
E-buffered RC step in the transient, and the ledger: every controlled
source's power is model-supplied and the ledger balances (the oracle fails
the analysis otherwise). A sign or transposition error in any of the four
stamps changes the golden's voltages.

GPU (marked gpu): operating points with element reports and ledgers, DC
sweeps, AC sweeps and transients of such circuits, batched with perturbed
gains, bit-identical to the oracle (linear circuits: the same operations in
the same order).
"""
import math

import numpy as np
import pytest

from kronspice_amd import api, circuit as kc, netlist
from oracle import oracle

ALL4 = """all four controlled sources with capacitive loads
V1 a 0 DC 2 AC 1 SIN(0 1 1k)
R1 a 0 1k
R2 a m 1k
VS m 0 0
E1 e 0 a 0 3
RE e x 1k
CE x 0 100n
G1 0 g a 0 2m
RG g 0 500
CG g 0 200n
H1 h 0 VS 500
RH h y 1k
CH y 0 50n
F1 0 f VS 3
RF f 0 100
CF f 0 1u
.end
"""


def flat(src, step=0.0, stop=0.0):
    c = netlist.parse(src).circuit
    els, p = c.flatten(step, stop)
    return c, els, p


def test_node_order_and_layout():
    """circuit.FromDeck registers an E/G element's control nodes before its
    output nodes (circuit/builder.go:409-412); E and H allocate a branch in
    element order, after V's (device/controlled.go:54-64)."""
    c = netlist.parse("t\nE1 out 0 in 0 2\nV1 in 0 1\nR1 out 0 1\n").circuit
    assert c.node_names == ["0", "in", "out"]
    lay = c.layout()
    assert lay.n == 4 and lay.branch_of == {0: 2, 1: 3}
    c, els, p = flat(ALL4)
    k = {e.name: i for i, e in enumerate(c.elements)}
    assert list(els[k["H1"]]["node"][:3]) == [c.node("h"), 0, k["VS"]]
    assert list(els[k["G1"]]["node"]) == [0, c.node("g"), c.node("a"), 0]


def test_controlling_source_must_exist():
    c = netlist.parse("t\nV1 a 0 1\nR1 a 0 1\nF1 0 a VX 2\n").circuit
    with pytest.raises(ValueError, match="controlling source"):
        c.flatten()
    # a controlling element without a branch of its own (a resistor)
    c, els, p = flat("t\nV1 a 0 1\nR1 a 0 1\nF1 0 a V1 2\n")
    els = els.copy()
    els[2]["node"][2] = 1  # R1
    st, _, _ = oracle.op(c.num_nodes, els, p)
    assert st == 10  # ErrUnsupported


def test_op_closed_form_and_ledger():
    c, els, p = flat(ALL4)
    st, x, info = oracle.op(c.num_nodes, els, p)
    assert st == 0
    v = {n: x[c.node(n) - 1] for n in ("a", "e", "g", "h", "f")}
    i_vs = 2.0 / 1000  # VS carries V1's current through R2 (m is at 0 V)
    assert math.isclose(v["e"], 6.0, rel_tol=1e-14)
    assert math.isclose(v["g"], 2e-3 * 2.0 * 500, rel_tol=1e-14)
    assert math.isclose(v["h"], 500 * i_vs, rel_tol=1e-14)
    assert math.isclose(v["f"], 3 * i_vs * 100, rel_tol=1e-14)
    st, _, rep, led = oracle.op_report(c.num_nodes, els, p)
    assert st == 0
    for name in ("E1", "F1", "G1", "H1"):
        r = rep[c.element_index(name)]
        assert r[5] == -r[2]  # model-supplied = -(absorbed)
    assert rep[c.element_index("F1")][2] < 0 and rep[c.element_index("G1")][2] < 0  # they deliver power
    assert rep[c.element_index("E1")][2] == 0  # no DC load behind the capacitor
    assert abs(led[6]) <= 1e-15 * led[7]  # residual against scale


@pytest.mark.parametrize("f", [10.0, 1e3, 1e5])
def test_ac_closed_forms(f):
    c, els, p = flat(ALL4)
    st, X, _ = oracle.ac(c.num_nodes, els, p, [f])
    assert st == 0
    w = 2 * math.pi * f
    va = 1.0  # AC 1 on V1; R1, R2 and VS leave a at the source
    i_vs = va / 1000
    want = {  # E and H force their outputs; RC dividers behind them, RC loads on G and F
        "e": 3 * va,
        "x": 3 * va / (1 + 1j * w * 1e3 * 100e-9),
        "g": 2e-3 * va * 500 / (1 + 1j * w * 500 * 200e-9),
        "h": 500 * i_vs,
        "y": 500 * i_vs / (1 + 1j * w * 1e3 * 50e-9),
        "f": 3 * i_vs * 100 / (1 + 1j * w * 100 * 1e-6),
    }
    for n, z in want.items():
        got = X[0, c.node(n) - 1]
        assert abs(got - z) <= 1e-13 * abs(z), (n, got, z)


def test_tran_buffered_rc_step():
    """E buffers a step into an RC: v(out) = 2·(1 - exp(-t/RC)) after the step."""
    src = """buffered rc
V1 in 0 PULSE(0 1 0 1n 1n 1 2)
E1 b 0 in 0 2
R1 b out 1k
C1 out 0 1u
.tran 10u 5m
.end
"""
    c = netlist.parse(src).circuit
    els, p = c.flatten(10e-6, 5e-3)
    spec = oracle.KhTranSpec(step=10e-6, stop=5e-3, start=0.0, max_step=0.0, uic=0)
    st, x, info = oracle.tran(c.num_nodes, els, p, spec, cap=200000)
    assert st == 0
    t = info["times"]
    v = info["xs"][:, c.node("out") - 1]
    m = t > 1e-5
    want = 2 * (1 - np.exp(-(t[m] - 1e-9) / 1e-3))
    assert np.max(np.abs(v[m] - want)) <= 2e-3 * 2


# ------------------------------------------------------------------ GPU
def perturbed(c, els, p, B, seed):
    rng = np.random.default_rng(seed)
    P = np.repeat(p[:, None], B, axis=1).copy()
    for k, e in enumerate(c.elements):
        if e.kind in (kc.E, kc.F, kc.G, kc.H):
            s = np.exp(rng.uniform(np.log(1 / 3), np.log(3), B))
            s[0] = 1.0
            P[els[k]["poff"]] = p[els[k]["poff"]] * s
    return P


@pytest.mark.gpu
def test_gpu_op_reports_bitexact():
    c, els, p = flat(ALL4)
    B = 64
    P = perturbed(c, els, p, B, 3)
    t = api.Topology(c.num_nodes, els, p)
    r = api.op(t, P, report=True)
    x = r.x.cpu().numpy()
    assert (r.status.cpu().numpy() == 0).all()
    rep = r.report.cpu().numpy()          # [B, nel, 8]
    led = r.report_ledger.cpu().numpy()   # [B, 10]
    for d in range(0, B, 7):
        st, xo, _ = oracle.op(c.num_nodes, els, P[:, d].copy())
        assert st == 0 and np.array_equal(x[:, d], xo), d
        st, _, ro, lo = oracle.op_report(c.num_nodes, els, P[:, d].copy())
        assert np.array_equal(rep[d], ro), d
        assert np.array_equal(led[d], lo), d


@pytest.mark.gpu
def test_gpu_dc_sweep_bitexact():
    """V1 swept from -2 V to 2 V: every point of every circuit (one gain draw
    each) equals the oracle's sweep of that circuit."""
    c, els, p = flat(ALL4)
    B = 8
    P = perturbed(c, els, p, B, 9)
    sweeps = [(c.element_index("V1"), -2.0, 2.0, 0.25)]
    t = api.Topology(c.num_nodes, els, p)
    r = api.dc(t, P, sweeps)
    x, st = r.x.cpu().numpy(), r.status.cpu().numpy()
    assert (st == 0).all()
    for d in range(B):
        e0, xo, ito, doneo, _ = oracle.dc(c.num_nodes, els, P[:, d].copy(), sweeps)
        assert e0 == 0 and np.array_equal(x[:, :, d], xo), d


@pytest.mark.gpu
def test_gpu_ac_bitexact():
    c, els, p = flat(ALL4)
    B = 16
    P = perturbed(c, els, p, B, 5)
    freqs = np.logspace(1, 6, 21)
    t = api.Topology(c.num_nodes, els, p)
    for backend in (0, 2):
        r = api.ac(t, P, freqs, opts=api.Options(backend=backend))
        assert (r.status.cpu().numpy() == 0).all()
        X = r.x.cpu().numpy()
        for d in range(B):
            e, xo, _ = oracle.ac(c.num_nodes, els, P[:, d].copy(), freqs, opts=oracle.options(backend=backend))
            assert e == 0 and np.array_equal(X[:, :, d], xo), (backend, d)


@pytest.mark.gpu
def test_gpu_tran_bitexact():
    c, els, p = flat(ALL4, 5e-6, 2e-3)
    B = 32
    P = perturbed(c, els, p, B, 7)
    t = api.Topology(c.num_nodes, els, p)
    spec = api.TranSpec(step=5e-6, stop=2e-3)
    r = api.tran(t, P, spec)
    x = r.x.cpu().numpy()
    acc = r.accepted.cpu().numpy()
    assert (r.status.cpu().numpy() == 0).all()
    ospec = oracle.KhTranSpec(step=5e-6, stop=2e-3, start=0.0, max_step=0.0, uic=0)
    for d in range(0, B, 5):
        st, xo, info = oracle.tran(c.num_nodes, els, P[:, d].copy(), ospec, cap=200000)
        assert st == 0 and np.array_equal(x[:, d], xo) and acc[d] == info["accepted"], d
