"""Mutual inductance, SPICE K (device/mutual.go:13-157, without the
spin-connection terms).

Oracle pins (CPU): the reference's mutual_ac and mutual_tran goldens and their
ngspice twins (run with the other goldens in tests/test_oracle.py), and what
the circuit equations fix — the loop equations of a resistively loaded
transformer solved by numpy at several frequencies (a wrong sign, a missing
factor or a transposed coupling stamp changes them), the stored energy M·i1·i2
at a DC point with currents in both coils (½L1·i1² + ½L2·i2² + M·i1·i2 being
the pair's energy), the ledger balance of a coupled transient, and
circuit.validateMutual's checks (circuit/circuit.go:804-830).

GPU (marked gpu): operating points with element reports, AC sweeps on both
backends and transients of coupled circuits, batched with perturbed couplings,
bit-identical to the oracle (linear circuits: the same operations in the same
order).
"""
import math

import numpy as np
import pytest

from kronspice_amd import api, circuit as kc, netlist
from oracle import oracle

TRANSFORMER = """resistively loaded transformer
V1 in 0 DC 0 AC 1 PULSE(0 1 0 1n 1n 1 2)
R1 in p 10
L1 p 0 1m
L2 s 0 4m
K1 L1 L2 0.8
R2 s 0 100
.end
"""

DC_CURRENTS = """currents in both coils at DC
V1 in 0 DC 1
R1 in p 10
L1 p 0 1m
I2 0 s DC 2m
L2 s 0 4m
K1 L1 L2 0.5
R2 s 0 100
.end
"""


def flat(src, step=0.0, stop=0.0):
    c = netlist.parse(src).circuit
    els, p = c.flatten(step, stop)
    return c, els, p


def test_element_table_and_m():
    """K's node[0], node[1] hold the coupled inductors' indices; its parameter
    is M = k·sqrt(L1·L2) (device/device.go:407-416)."""
    c, els, p = flat(TRANSFORMER)
    k = c.element_index("K1")
    assert list(els[k]["node"][:2]) == [c.element_index("L1"), c.element_index("L2")]
    assert p[els[k]["poff"]] == 0.8 * math.sqrt(1e-3 * 4e-3)
    assert c.num_nodes == 4  # the K element adds no node


@pytest.mark.parametrize("bad, msg", [
    ("K1 L1 L2 0", "coupling coefficient"),
    ("K1 L1 L2 1.5", "coupling coefficient"),
    ("K1 L1 L1 0.5", "with itself"),
    ("K1 L1 R2 0.5", "not a linear inductor"),
    ("K1 L1 L2 0.5\nK2 L2 L1 0.3", "already coupled"),
])
def test_validate_mutual(bad, msg):
    src = "t\nV1 in 0 1\nR1 in p 10\nL1 p 0 1m\nL2 s 0 4m\nR2 s 0 100\n" + bad + "\n.end\n"
    with pytest.raises(ValueError, match=msg):
        netlist.parse(src).circuit.flatten()


@pytest.mark.parametrize("f", [50.0, 1e3, 1e5])
def test_ac_loop_equations(f):
    """v(p) = jwL1·i1 + jwM·i2, v(s) = jwM·i1 + jwL2·i2 with i1 = (1 - v(p))/R1
    and i2 = -v(s)/R2: the 2x2 loop system solved directly."""
    c, els, p = flat(TRANSFORMER)
    st, X, _ = oracle.ac(c.num_nodes, els, p, [f])
    assert st == 0
    w = 2 * math.pi * f
    L1, L2, R1, R2 = 1e-3, 4e-3, 10.0, 100.0
    M = 0.8 * math.sqrt(L1 * L2)
    A = np.array([[R1 + 1j * w * L1, 1j * w * M], [1j * w * M, R2 + 1j * w * L2]])
    i1, i2 = np.linalg.solve(A, np.array([1.0, 0.0]))
    vp = 1j * w * L1 * i1 + 1j * w * M * i2
    vs = -R2 * i2
    for n, z in (("p", vp), ("s", vs)):
        got = X[0, c.node(n) - 1]
        assert abs(got - z) <= 1e-12 * abs(z), (n, got, z)


def test_dc_stored_energy_and_report():
    """At DC the coils are shorts: i1 = 1 V / 10 ohm, i2 = 2 mA from I2. The K
    entry stores M·i1·i2 and has no voltage, current, power or port."""
    c, els, p = flat(DC_CURRENTS)
    st, x, rep, led = oracle.op_report(c.num_nodes, els, p)
    assert st == 0
    lay = c.layout()
    i1 = x[lay.branch_of[c.element_index("L1")]]
    i2 = x[lay.branch_of[c.element_index("L2")]]
    assert math.isclose(i1, 0.1, rel_tol=1e-12) and math.isclose(i2, 2e-3, rel_tol=1e-12)
    M = 0.5 * math.sqrt(1e-3 * 4e-3)
    r = rep[c.element_index("K1")]
    assert r[4] == M * i1 * i2
    assert r[0] == 0 and r[1] == 0 and r[2] == 0 and r[3] == 0 and r[5] == 0 and r[6] == 0
    w = 0.5 * 1e-3 * i1 ** 2 + 0.5 * 4e-3 * i2 ** 2 + r[4]
    assert math.isclose(rep[c.element_index("L1")][4] + rep[c.element_index("L2")][4] + r[4], w, rel_tol=1e-14)
    assert abs(led[6]) <= 1e-15 * led[7]


@pytest.mark.parametrize("kval", [0.3, 0.8, 1.0])
def test_tran_pulse_coupling(kval):
    """A 1 V step into the primary: the analysis completes with its ledger
    balanced at every point (the oracle fails it otherwise), and the secondary
    current's sign is the coupling's: i2 flows against i1's rise."""
    src = TRANSFORMER.replace("K1 L1 L2 0.8", f"K1 L1 L2 {kval}")
    c, els, p = flat(src, 2e-6, 2e-4)
    spec = oracle.KhTranSpec(step=2e-6, stop=2e-4, start=0.0, max_step=0.0, uic=0)
    st, x, info = oracle.tran(c.num_nodes, els, p, spec, cap=200000)
    assert st == 0
    lay = c.layout()
    xs = info["xs"]
    i1 = xs[:, lay.branch_of[c.element_index("L1")]]
    vs = xs[:, c.node("s") - 1]
    early = (info["times"] > 1e-6) & (info["times"] < 2e-5)
    assert (i1[early] > 0).all() and (vs[early] > 0).all()  # v(s) = -R2·i2 > 0: i2 < 0


# ------------------------------------------------------------------ GPU
def perturbed(c, els, p, B, seed):
    rng = np.random.default_rng(seed)
    P = np.repeat(p[:, None], B, axis=1).copy()
    for k, e in enumerate(c.elements):
        if e.kind == kc.K:
            s = rng.uniform(0.5, 1.0, B)
            s[0] = 1.0
            P[els[k]["poff"]] = p[els[k]["poff"]] * s
        elif e.kind == kc.R:
            s = np.exp(rng.uniform(np.log(0.5), np.log(2), B))
            s[0] = 1.0
            P[els[k]["poff"]] = p[els[k]["poff"]] * s
            P[els[k]["poff"] + 1] = 1 / P[els[k]["poff"]]
    return P


@pytest.mark.gpu
def test_gpu_op_reports_bitexact():
    c, els, p = flat(DC_CURRENTS)
    B = 64
    P = perturbed(c, els, p, B, 3)
    t = api.Topology(c.num_nodes, els, p)
    r = api.op(t, P, report=True)
    x = r.x.cpu().numpy()
    assert (r.status.cpu().numpy() == 0).all()
    rep = r.report.cpu().numpy()
    led = r.report_ledger.cpu().numpy()
    for d in range(0, B, 7):
        st, xo, _ = oracle.op(c.num_nodes, els, P[:, d].copy())
        assert st == 0 and np.array_equal(x[:, d], xo), d
        st, _, ro, lo = oracle.op_report(c.num_nodes, els, P[:, d].copy())
        assert np.array_equal(rep[d], ro), d
        assert np.array_equal(led[d], lo), d


@pytest.mark.gpu
def test_gpu_ac_bitexact():
    c, els, p = flat(TRANSFORMER)
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
    c, els, p = flat(TRANSFORMER, 2e-6, 2e-4)
    B = 32
    P = perturbed(c, els, p, B, 7)
    t = api.Topology(c.num_nodes, els, p)
    spec = api.TranSpec(step=2e-6, stop=2e-4)
    r = api.tran(t, P, spec)
    x = r.x.cpu().numpy()
    acc, rej = r.accepted.cpu().numpy(), r.rejected.cpu().numpy()
    assert (r.status.cpu().numpy() == 0).all()
    ospec = oracle.KhTranSpec(step=2e-6, stop=2e-4, start=0.0, max_step=0.0, uic=0)
    for d in range(0, B, 5):
        st, xo, info = oracle.tran(c.num_nodes, els, P[:, d].copy(), ospec, cap=200000)
        assert st == 0 and acc[d] == info["accepted"] and rej[d] == info["rejected"], d
        assert np.array_equal(x[:, d], xo), d


def test_subckt_coupled_pair():
    """A K element inside a .subckt couples that instance's inductors
    (netlist/preprocess.go:649-651): two instances give two independent
    transformers with the top-level deck's answer."""
    sub = """two transformers
.subckt xf in out
R1 in p 10
L1 p 0 1m
L2 out 0 4m
K1 L1 L2 0.8
R2 out 0 100
.ends
V1 a 0 DC 0 AC 1
X1 a s1 xf
X2 a s2 xf
.end
"""
    c = netlist.parse(sub).circuit
    els, p = c.flatten()
    st, X, _ = oracle.ac(c.num_nodes, els, p, [1e3])
    assert st == 0
    ct, et, pt = flat(TRANSFORMER)
    st, Xt, _ = oracle.ac(ct.num_nodes, et, pt, [1e3])
    want = Xt[0, ct.node("s") - 1]
    for n in ("s1", "s2"):
        assert abs(X[0, c.node(n) - 1] - want) <= 1e-13 * abs(want), n
