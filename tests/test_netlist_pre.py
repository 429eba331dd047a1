"""Netlist preprocessing (kronspice_amd/netlist_pre.py): .param, .global,
.subckt / X expansion and expressions, against the known answers of the
reference's own tests (netlist/preprocess_test.go:17-216) restated here, and
the two golden decks that need it (param_rc_ac, subckt_ladder: the golden
runner's oracle checks cover their analyses)."""
import math

import pytest

from kronspice_amd import netlist
from kronspice_amd.netlist_pre import ParamError, SubcktError, evaluate, expand

EXPR_CASES = [
    ("1+2*3", 7), ("(1+2)*3", 9), ("2^3^2", 512), ("2**3", 8), ("-2^2", -4),
    ("10k/2", 5000), ("1meg*1u", 1), ("4.7kohm", 4700), ("sqrt(16)+abs(-1)", 5),
    ("max(1, 3) - min(2,5)", 1), ("2*pi", 2 * math.pi), ("log10(1000)", 3),
    ("ln(exp(2))", 2), ("1 < 2", 1), ("1 >= 2", 0), ("1 == 1 && 0 || 1", 1),
    ("!0", 1), ("3 != 3", 0), ("1 ? 10 : 20", 10), ("0 ? 10 : 2>1 ? 30 : 40", 30),
    ("7 % 4", 3), ("pwr(-2, 2)", -4), ("nint(2.5)", 2), (".5e1", 5), ("sgn(-3)", -1),
]


@pytest.mark.parametrize("expr,want", EXPR_CASES)
def test_eval_expr_known_answers(expr, want):
    got = evaluate(expr)
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (expr, got, want)


@pytest.mark.parametrize("bad", ["", "1+", "(1", "1/0", "foo", "sqrt(1,2)", "nope(1)", "1 2", "ln(0)",
                                 "sqrt(-1)", "(" * 500 + "1" + ")" * 500])
def test_eval_expr_errors(bad):
    with pytest.raises(ParamError):
        evaluate(bad)


PARAM_DECK = """params
.param rtop = {2*rbot} rbot=1k
.param vin='5*gain', gain=2
V1 in 0 DC {vin}
R1 in out {rtop}
R2 out 0 'rbot'
C1 out 0 {1/(2*pi*rbot*1k)} IC={vin/2}
.tran {1u} {10*1u}
.end
"""


def test_param_deck():
    d = netlist.parse(PARAM_DECK)
    vals = {e.name: e.value for e in d.circuit.elements}
    want = {"V1": 10.0, "R1": 2000.0, "R2": 1000.0, "C1": 1 / (2 * math.pi * 1e6)}
    for k, w in want.items():
        assert abs(vals[k] - w) <= 1e-12 * w, (k, vals[k], w)
    c1 = [e for e in d.circuit.elements if e.name == "C1"][0]
    assert c1.ic == 5.0
    a = d.analyses[0]
    assert a.args["step"] == 1e-6 and abs(a.args["stop"] - 1e-5) <= 1e-20
    lines = [ln.strip() for ln in PARAM_DECK.splitlines()[1:] if ln.strip()]
    _, params = expand(lines)
    assert params["rtop"] == 2000 and params["gain"] == 2 and len(params) == 4


@pytest.mark.parametrize("bad", [".param a={b} b={a}\nR1 1 0 {a}", ".param a=1 a=2\nR1 1 0 {a}", "R1 1 0 {undefined}",
                                 ".param x={1/0}\nR1 1 0 1", ".param =3", "R1 1 0 {1", ".param sqrt=2"])
def test_param_errors(bad):
    with pytest.raises(ParamError):
        netlist.parse("t\n" + bad + "\n")


# The reference's subcircuit deck (netlist/preprocess_test.go:88-115): naming,
# scoping, defaults, nested definitions, a controlling source named inside an
# instance, local models.
SUBCKT_DECK = """subcircuits
.param rload=4k
.subckt divider top mid params: rt=1k rb={rt}
R1 top mid {rt}
R2 mid 0 {rb}
.ends divider
.subckt stage in out params: g=2
.subckt buf a b
E1 b 0 a 0 1
.ends
Vs in sense 0
X1 sense n divider rt={1k*g}
X2 n out buf
F1 0 out Vs {g}
D1 out 0 dloc
.model dloc D(IS=1e-14)
.ends
V1 vin 0 DC 10
XA vin a divider
XB vin b divider rt=2k rb={rload}
XS a s stage params: g=3
.op
.end
"""


def test_subckt_expansion():
    d = netlist.parse(SUBCKT_DECK)
    c = d.circuit
    by = {e.name: e for e in c.elements}
    name = lambda k: c.node_names[k]

    def check(nm, value, *nodes):
        e = by[nm]
        if value is not None:
            assert e.value == value, (nm, e.value, value)
        assert tuple(name(k) for k in e.nodes[:len(nodes)]) == nodes, (nm, [name(k) for k in e.nodes])

    check("R.XA.R1", 1000.0, "vin", "a")
    check("R.XA.R2", 1000.0, "a", "0")
    check("R.XB.R1", 2000.0, "vin", "b")
    check("R.XB.R2", 4000.0, "b", "0")
    check("V.XS.Vs", 0.0, "a", "xs.sense")
    check("R.XS.X1.R1", 3000.0, "xs.sense", "xs.n")
    check("R.XS.X1.R2", 3000.0, "xs.n", "0")  # rb defaults to rt
    check("E.XS.X2.E1", 1.0, "s", "0")
    assert tuple(name(k) for k in by["E.XS.X2.E1"].ctrl_nodes) == ("xs.n", "0")
    check("F.XS.F1", 3.0, "0", "s")
    assert by["F.XS.F1"].ctrl_source == "v.xs.vs"
    check("D.XS.D1", None, "s", "0")
    assert len(c.elements) == 11
    lines = [ln.strip() for ln in SUBCKT_DECK.splitlines()[1:] if ln.strip()]
    flat, _ = expand(lines)
    assert any(ln.lower().startswith(".model xs.dloc") for ln in flat)
    assert any(ln.startswith("D.XS.D1") and ln.split()[3] == "xs.dloc" for ln in flat)


@pytest.mark.parametrize("src,err", [
    (".subckt r2 a b\nR1 a b 1\n", SubcktError),
    (".ends\n", SubcktError),
    (".subckt r2 a b params: r=1\nR1 a b {r}\n.ends\nX1 1 r2\n", SubcktError),
    (".subckt r2 a b params: r=1\nR1 a b {r}\n.ends\nX1 1 0 nosuch\n", SubcktError),
    (".subckt r2 a b params: r=1\nR1 a b {r}\n.ends\nX1 1 0 r2 q=3\n", ParamError),
    (".subckt loop a\nX1 a loop\n.ends\nX1 1 loop\n", SubcktError),
    (".subckt s a\n.tran 1 2\n.ends\nX1 1 s\n", SubcktError),
    (".subckt s a a\n.ends\n", SubcktError),
    (".subckt r2 a b params: r=1\nR1 a b {r}\n.ends\n" * 2, SubcktError),
    (".subckt s a\n.ends t\n", SubcktError),
])
def test_subckt_errors(src, err):
    with pytest.raises(err):
        netlist.parse("t\n" + src)


def test_subckt_global_and_scope():
    d = netlist.parse("""scopes
.global vdd
.param r=10
.subckt cell a
.param r2={2*r}
R1 a vdd {r2}
R2 a mid {r}
.ends
X1 n cell
""")
    c = d.circuit
    e1, e2 = c.elements[0], c.elements[1]
    assert e1.value == 20 and c.node_names[e1.nodes[1]] == "vdd"
    assert e2.value == 10 and c.node_names[e2.nodes[1]] == "x1.mid"


def test_golden_decks_flatten():
    from tests import golden_runner as gr
    d = gr.load_deck("param_rc_ac")
    cv = [e for e in d.circuit.elements if e.name == "C1"][0].value
    assert cv == 1 / (2 * math.pi * 1e3 * 1e3)  # the same operations in the same order
    d = gr.load_deck("subckt_ladder")
    assert [e.name for e in d.circuit.elements] == ["V1", "R.XA.X1.R1", "R.XA.X1.R2", "R.XA.X2.R1", "R.XA.X2.R2",
                                                     "R.XB.R1", "R.XB.R2", "RL"]
