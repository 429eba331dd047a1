"""output.Result's element entries and ledger (output/result.go:15-109), as
analysis's Sim.result builds them (analysis/op.go:361-404) and em.BuildLedger
totals them (em/energy.go:89-119).

Oracle pins (CPU):
  * the reference's own known answer (cmd/kronspice/main_test.go:62-95, the
    divider deck): v(out) = 5 V, three elements, inflow = dissipation = 50 mW,
    relative residual 0; the source's power enters from the exterior, the
    resistors' leaves as heat;
  * stored energies where mathematics fixes them: ½·C·v² and ½·L·i²; a
    diode's co-energy ∫₀^vd v·C(v) dv against an independent numerical
    integral of the SPICE junction capacitance (depletion and diffusion, each
    bias region: forward, below FC·VJ and above, reverse, breakdown); MOSFET
    and BJT junctions with grading coefficient 0, whose charges are linear
    (q = C·v), so the quadrature of the reference must give ½·C·v² exactly to
    round-off;
  * the ledger's bookkeeping: residual = inflow + model-supplied − outflow −
    dissipation − storage rate, scale = Σ|terms|, and the element powers sum
    to zero (Tellegen) on every golden operating point.

GPU parity: the report kernels (kh_report.h, k_report / k_report_ledger)
against the oracle on the golden decks' operating points and on batches.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from kronspice_amd import netlist
from oracle import oracle
from tests import golden_runner as gr

V, I, P, S, W, MS, EXT, CH = range(8)
INFLOW, OUTFLOW, DISS, STOR, SUP, HYP, RES, SCALE, RO, REL = range(10)

DIVIDER = """Divider
V1 in 0 DC 10
R1 in out 1k
R2 out 0 1k
.op
.end
"""


def deck_report(text, polish=False, opts=None):
    c = netlist.parse(text).circuit
    els, p = c.flatten()
    e, x, rep, led = oracle.op_report(c.num_nodes, els, p, opts=opts, polish=polish)
    return c, els, p, e, x, rep, led


def test_divider_known_answer():
    c, els, p, e, x, rep, led = deck_report(DIVIDER)
    assert e == 0 and len(rep) == 3
    assert abs(x[c.node("out") - 1] - 5) <= 1e-9
    assert abs(led[INFLOW] - 0.05) <= 1e-9 and abs(led[DISS] - 0.05) <= 1e-9 and led[REL] <= 1e-9
    assert led[OUTFLOW] == 0 and led[STOR] == 0 and led[SUP] == 0 and led[HYP] == 0
    v1, r1, r2 = rep
    assert v1[CH] == 0 and v1[EXT] == pytest.approx(0.05, abs=1e-15) and v1[P] == pytest.approx(-0.05, abs=1e-15)
    assert v1[I] == pytest.approx(-0.005, abs=1e-15)  # the branch current, n+ -> n- through the source
    for r in (r1, r2):
        assert r[CH] == 2 and r[EXT] == -r[P] and r[P] == pytest.approx(0.025, abs=1e-15)
        assert r[V] == pytest.approx(5, abs=1e-12) and r[I] == pytest.approx(0.005, abs=1e-15)


def test_capacitor_and_inductor_energies():
    text = """energies
V1 a 0 DC 10
C1 a 0 2u
R1 a b 1k
L1 b 0 3m
.op
.end
"""
    c, els, p, e, x, rep, led = deck_report(text)
    assert e == 0
    cap, ind = rep[1], rep[3]
    assert cap[W] == 0.5 * 2e-6 * 10.0 * 10.0
    i = x[-1]  # the inductor's branch current (its branch is the last unknown)
    assert abs(i - 0.01) <= 1e-12
    assert ind[W] == 0.5 * 3e-3 * i * i
    assert cap[CH] == -1 and ind[CH] == -1 and cap[EXT] == 0 and ind[EXT] == 0


def negative_resistor_deck():
    return """neg
V1 a 0 DC 1
R1 a b 50
Rn b 0 -200
.op
.end
"""


def test_negative_resistance_is_model_supplied():
    c, els, p, e, x, rep, led = deck_report(negative_resistor_deck())
    assert e == 0
    rn = rep[2]
    assert rn[CH] == -1 and rn[EXT] == 0 and rn[MS] == -rn[P] and rn[MS] > 0
    assert led[SUP] == rn[MS]
    assert led[RES] == led[INFLOW] + led[SUP] - led[OUTFLOW] - led[DISS] - led[STOR]


# ---- diode co-energy --------------------------------------------------------------
def diode_deck(vd, model):
    return f"""diode
V1 a 0 DC {vd!r}
D1 a 0 dm
.model dm D({model})
.op
.end
"""


def diode_params(c, els, p):
    k = [i for i, e in enumerate(c.elements) if e.name.lower() == "d1"][0]
    q = p[els[k]["poff"]:]
    return dict(IS=q[0], CJO=q[3], VJ=q[4], M=q[5], TT=q[6], FC=q[9], hasBV=q[10] != 0, vte=q[11], xbv=q[13])


def junction_current(v, d):
    """SPICE diode junction current by region (device/diode.go:148-162)."""
    IS, vte = d["IS"], d["vte"]
    if v >= -3 * vte:
        return IS * math.expm1(min(v / vte, 700))
    if not d["hasBV"] or v >= -d["xbv"]:
        a = 3 * vte / (v * math.e)
        return -IS * (1 + a ** 3)
    return -IS * math.exp(min(-(d["xbv"] + v) / vte, 700))


def depletion_capacitance(v, d):
    """dq/dv of the depletion charge: CJO·(1−v/VJ)^−M below FC·VJ, the linear
    extension above (device/diode.go:191-211)."""
    CJO, VJ, M, FC = d["CJO"], d["VJ"], d["M"], d["FC"]
    if CJO == 0:
        return 0.0
    if v < FC * VJ:
        return CJO * (1 - v / VJ) ** (-M)
    f2 = (1 - FC) ** (1 + M)
    return CJO / f2 * (1 - FC * (1 + M) + M * v / VJ)


def numeric_coenergy(vd, d):
    """∫₀^vd v·C(v) dv with C = TT·did/dv + Cj: the diffusion part integrated
    by parts (v·TT·id − TT·∫id), everything by adaptive quadrature."""
    brk = sorted({0.0, vd, d["FC"] * d["VJ"], -3 * d["vte"]} | ({-d["xbv"]} if d["hasBV"] else set()))
    brk = [b for b in brk if min(0, vd) <= b <= max(0, vd)]

    def quad(f):
        tot = 0.0
        for a, b in zip(brk[:-1], brk[1:]):
            tot += integrate.quad(f, a, b, epsabs=0, epsrel=1e-13, limit=400)[0]
        return tot if vd >= 0 else -tot

    wj = quad(lambda v: v * depletion_capacitance(v, d))
    wd = d["TT"] * (vd * junction_current(vd, d) - quad(lambda v: junction_current(v, d)))
    return wj + wd


@pytest.mark.parametrize("vd", [0.3, 0.55, 0.72, -0.01, -0.5, -3.0])
@pytest.mark.parametrize("model", ["IS=1e-14 CJO=2p VJ=0.7 M=0.45 TT=10n",
                                   "IS=2e-15 N=1.5 CJO=5p VJ=0.9 M=0.33 FC=0.4 TT=3n",
                                   "IS=1e-14 CJO=1p M=0 TT=0"])
def test_diode_coenergy_against_numerical_integral(vd, model):
    c, els, p, e, x, rep, led = deck_report(diode_deck(vd, model))
    assert e == 0
    d = diode_params(c, els, p)
    got, want = rep[1][W], numeric_coenergy(vd, d)
    assert abs(got - want) <= 1e-9 * abs(want) + 1e-30, (got, want)
    if d["M"] == 0 and d["TT"] == 0:  # linear junction charge: exactly ½·CJO·v²
        assert got == pytest.approx(0.5 * d["CJO"] * vd * vd, rel=1e-12, abs=0)  # cancellation in v·q − ∫q


@pytest.mark.parametrize("vs", [-4.9, -5.3, -8.0, -20.0])
def test_diode_coenergy_reverse_and_breakdown(vs):
    """Reverse bias through 1 kΩ: below breakdown, and in it (the junction
    voltage then sits just past −XBV)."""
    model = "IS=1e-14 CJO=2p VJ=0.7 M=0.5 TT=5n BV=5 IBV=1m"
    text = f"""rev
V1 a 0 DC {vs!r}
R1 a b 1k
D1 b 0 dm
.model dm D({model})
.op
.end
"""
    c, els, p, e, x, rep, led = deck_report(text)
    assert e == 0
    d = diode_params(c, els, p)
    vd = x[c.node("b") - 1]
    if vs < -6:
        assert vd < -d["xbv"]  # in breakdown
    got, want = rep[2][W], numeric_coenergy(vd, d)
    assert abs(got - want) <= 1e-9 * abs(want), (got, want)


# ---- linear junction charges: the quadratures reduce to ½·C·v² -------------------------
def test_mosfet_linear_junction_and_overlap_energy():
    text = """mos
VD d 0 DC 2.5
VG g 0 DC 1.2
VB b 0 DC -0.7
M1 d g 0 b nm W=10u L=1u
.model nm NMOS(LEVEL=1 VTO=0.6 KP=100u CBD=20f CBS=30f MJ=0 MJSW=0 PB=0.8 CGSO=1n CGDO=2n CGBO=3n)
.op
.end
"""
    c, els, p, e, x, rep, led = deck_report(text)
    assert e == 0
    k = [i for i, el in enumerate(c.elements) if el.name.lower() == "m1"][0]
    q = p[els[k]["poff"]:]
    czbd, czbdsw, czbs, czbssw, cgso, cgdo, cgbo, ox = q[13], q[14], q[15], q[16], q[27], q[28], q[29], q[30]
    assert ox == 0  # no TOX: no Meyer capacitances, only the overlaps
    vg, vd, vb = x[c.node("g") - 1], x[c.node("d") - 1], x[c.node("b") - 1]
    vgs, vgd, vgb, vbs, vbd = vg, vg - vd, vg - vb, vb, vb - vd
    want = 0.5 * (czbs + czbssw) * vbs ** 2 + 0.5 * (czbd + czbdsw) * vbd ** 2 + \
        0.5 * cgso * vgs ** 2 + 0.5 * cgdo * vgd ** 2 + 0.5 * cgbo * vgb ** 2
    assert want > 0
    assert rep[k][W] == pytest.approx(want, rel=1e-13, abs=0)


def test_bjt_linear_junction_energy():
    text = """bjt
VC c 0 DC 5
VB b 0 DC 0.65
VS s 0 DC -2
Q1 c b 0 s qn
.model qn NPN(IS=1e-16 BF=100 CJE=2p VJE=0.75 MJE=0 CJC=1p VJC=0.75 MJC=0 XCJC=0.6 CJS=3p VJS=0.75 MJS=0 TF=0 TR=0)
.op
.end
"""
    c, els, p, e, x, rep, led = deck_report(text)
    assert e == 0
    k = [i for i, el in enumerate(c.elements) if el.name.lower() == "q1"][0]
    q = p[els[k]["poff"]:]
    CJE, CJC, XCJC, CJS = q[19], q[26], q[29], q[31]
    vc, vb, vs = x[c.node("c") - 1], x[c.node("b") - 1], x[c.node("s") - 1]
    vbe, vbc, vcs = vb, vb - vc, vc - vs
    want = 0.5 * CJE * vbe ** 2 + 0.5 * CJC * XCJC * vbc ** 2 + 0.5 * CJC * (1 - XCJC) * vbc ** 2 + \
        0.5 * CJS * vcs ** 2
    assert rep[k][W] == pytest.approx(want, rel=1e-13, abs=0)


# ---- the ledger's bookkeeping on every golden operating point ------------------------------
@pytest.mark.parametrize("name", [n for n in gr.names() if "tran" not in n and "_ac" not in n
                                  and n not in ("rc_step", "rl_step", "rlc_step", "rectifier_rc")])
def test_golden_ledgers_balance(name):
    c = gr.load_deck(name).circuit
    els, p = c.flatten()
    e, x, rep, led = oracle.op_report(c.num_nodes, els, p)
    assert e == 0
    assert led[RES] == led[INFLOW] + led[SUP] - led[OUTFLOW] - led[DISS] - led[STOR]
    # Tellegen: the absorbed powers of all elements sum to zero
    assert abs(rep[:, P].sum()) <= 1e-9 * max(led[SCALE], 1e-300) + 1e-18
    assert led[REL] <= 1e-9
    # per element: dissipated heat = absorbed − stored for the junction devices
    for k, el in enumerate(els):
        if el["kind"] in (6, 7, 8):
            assert rep[k][CH] == 2 and rep[k][EXT] == -(rep[k][P] - rep[k][S])


# ---- AC: element currents and the average-power ledger per frequency -------------------------------
def test_ac_rc_lowpass_currents_and_ledger_closed_forms():
    """RC low-pass driven by a 1 V AC source (cmd/kronspice/main_test.go:
    33-39): V_out = 1/(1 + jωRC); the resistor carries (1 − V_out)/R, the
    capacitor jωC·V_out, the source −(1 − V_out)/R; the source's average power
    ½·Re(I_R) enters, ½·|I_R|²·R is dissipated, the capacitor's reactive power
    is −½·ωC·|V_out|², and the residual is round-off (analysis/ac.go:148-173)."""
    text = """RC low-pass
V1 in 0 AC 1
R1 in out 1k
C1 out 0 159.15n
.ac dec 2 10 100k
.end
"""
    c = netlist.parse(text).circuit
    els, p = c.flatten()
    freqs = np.array([10.0, 1e3, 31622.776601683792, 1e5])
    e, X, mr, cur, led = oracle.ac(c.num_nodes, els, p, freqs, currents=True)
    assert e == 0
    R, Cc = 1e3, 159.15e-9
    for k, f in enumerate(freqs):
        w = 2 * math.pi * f
        vout = 1 / (1 + 1j * w * R * Cc)
        ir = (1 - vout) / R
        assert abs(cur[k][1] - ir) <= 1e-12 * abs(ir)
        assert abs(cur[k][2] - 1j * w * Cc * vout) <= 1e-12 * abs(ir)
        assert abs(cur[k][0] + ir) <= 1e-12 * abs(ir)
        inflow, outflow, diss, reactive, sup, res = led[k]
        # 1 − V_out cancels at low frequency: the solution's round-off is ~1e-12 relative there
        assert inflow == pytest.approx(0.5 * ir.real, rel=1e-9, abs=0)
        assert diss == pytest.approx(0.5 * abs(ir) ** 2 * R, rel=1e-9, abs=0)
        assert reactive == pytest.approx(-0.5 * w * Cc * abs(vout) ** 2, rel=1e-9, abs=0)
        assert outflow == 0 and sup == 0
        assert res == inflow + sup - outflow - diss and abs(res) <= 1e-10 * inflow
