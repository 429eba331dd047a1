"""The oracle's restatement of kron.JacobianTearing (oracle/tear.hpp), pinned
the way the reference pins JacobianTearing (kron/jacobian_test.go):

  * every golden netlist that runs here, under every single-element cut and
    the all-element cut (jacobianCuts), with every system of its analysis
    (OP, DC sweep, AC, transient) solved torn, matches its golden values
    within the golden tolerances and the untorn run within 1e-6 of scale;
    at least 40 torn runs, 30 of them split into several pieces, and at most
    one solve in 20 falls back to the untorn LU (:127-214);
  * the rectifier's RC ladder torn from it: the ladder's piece is factored
    once and reused, the waveform equals the untorn one to 1e-9 V, and with a
    reuse tolerance pieces are reused stale (:238-282);
  * the cut's errors (:284-308), a singular circuit stays singular (:300-307);
  * the 8-unknown chain torn in the middle: two pieces, no fallback (:339-373);
  * what the mathematics fixes: torn solutions of random clustered systems
    equal numpy's to round-off with no refinement step and no fallback (a
    wrong sign or index in the interconnection would need refinement or fall
    back), the equilibrated dense LU solves badly scaled systems, and a
    singular piece falls back to the untorn LU, whose result it returns.
"""
import math

import numpy as np
import pytest

from kronspice_amd import netlist, tear
from oracle import oracle
from tests import golden_runner as gr
from tests.test_oracle import OracleBackend


class TornBackend(OracleBackend):
    """The oracle's analyses with Options.Solver = a JacobianTearing of `cut`
    (every matrix the analysis makes is a torn one), recording the values of
    each call for the twin comparison."""

    def __init__(self, c, cut, reuse_tol=0.0):
        super().__init__(0)
        self.t = oracle.Tearing(tear.cut_pairs(c, cut), reuse_tol)
        self.vals = []

    def op(self, c):
        with self.t:
            st, x, cur = super().op(c)
        self.vals.append(np.asarray(x, float).ravel())
        return st, x, cur

    def tran(self, c, a):
        with self.t:
            r = super().tran(c, a)
        self.vals.append(np.concatenate([np.asarray(r[1]).ravel(), np.asarray(r[2]).ravel()]))
        return r

    def ac(self, c, freqs):
        with self.t:
            st, X = super().ac(c, freqs)
        self.vals.append(np.concatenate([np.asarray(X).real.ravel(), np.asarray(X).imag.ravel()]))
        return st, X


class Recording(OracleBackend):
    def __init__(self):
        super().__init__(0)
        self.vals = []

    def op(self, c):
        st, x, cur = super().op(c)
        self.vals.append(np.asarray(x, float).ravel())
        return st, x, cur

    def tran(self, c, a):
        r = super().tran(c, a)
        self.vals.append(np.concatenate([np.asarray(r[1]).ravel(), np.asarray(r[2]).ravel()]))
        return r

    def ac(self, c, freqs):
        st, X = super().ac(c, freqs)
        self.vals.append(np.concatenate([np.asarray(X).real.ravel(), np.asarray(X).imag.ravel()]))
        return st, X


def test_jacobian_tearing_goldens():
    runs = split = 0
    solves = fallbacks = factors = reuses = 0
    worst = 0.0
    for name in gr.names():
        c = gr.load_deck(name).circuit
        ref = Recording()
        ref_msgs = gr.check(name, ref)
        assert not ref_msgs, (name, ref_msgs[:3])
        for cut in tear.jacobian_cuts(c):
            b = TornBackend(gr.load_deck(name).circuit, cut)
            msgs = gr.check(name, b)
            assert not msgs, (name, cut, msgs[:3])
            runs += 1
            s = b.t.stats()
            split += s["max_pieces"] > 1
            solves += s["solves"]
            fallbacks += s["fallbacks"]
            factors += s["factors"]
            reuses += s["reuses"]
            if [g.shape for g in b.vals] != [w.shape for w in ref.vals]:
                continue  # a transient may take other steps; the golden check resampled it
            got, want = np.concatenate(b.vals), np.concatenate(ref.vals)  # a run's values (a DC sweep's points)
            scale = max(float(np.max(np.abs(want))), 1e-30)
            worst = max(worst, float(np.max(np.abs(got - want))) / scale)
    print(f"{runs} torn runs ({split} split); {solves} systems, {factors} pieces factored, {reuses} reused, "
          f"{fallbacks} fallbacks; worst twin difference {worst:.2g}")
    assert runs >= 40 and split >= 30, (runs, split)
    assert fallbacks * 20 <= solves, (fallbacks, solves)
    assert worst <= 1e-6, worst


DIODE_LADDER = """diode ladder: a fixed RC network torn from a rectifier
V1 in 0 SIN(0 5 1k)
R1 in a 100
D1 a b dmod
C1 b 0 1u
RL b 0 10k
RT b p 1k
R2 p q 1k
C2 q 0 100n
R3 q r 1k
C3 r 0 100n
R4 r 0 10k
.model dmod D(IS=1e-14 N=1.05 RS=0.5 CJO=2p)
.tran 10u 3m
.end
"""


def tran_of(c, t=None):
    b = OracleBackend(0)
    a = dict(step=10e-6, stop=3e-3, start=0.0, max_step=0.0, uic=False)
    if t is None:
        return b.tran(c, a)
    with t:
        return b.tran(c, a)


@pytest.mark.parametrize("tol", [0.0, 0.01])
def test_jacobian_reuse(tol):
    c = netlist.parse(DIODE_LADDER).circuit
    st0, t0, xs0, _, _ = tran_of(c)
    assert st0 == 0
    t = oracle.Tearing(tear.cut_pairs(c, ["RT"]), tol)
    st, t1, xs1, _, _ = tran_of(c, t)
    assert st == 0
    s = t.stats()
    assert s["max_pieces"] == 2 and s["fallbacks"] * 20 <= s["solves"], s
    assert s["reuses"] + s["stale"] >= s["solves"] // 2, s
    if tol > 0:
        assert s["stale"] > 0, s
    r = c.node("r") - 1
    assert len(t1) == len(t0)
    assert float(np.max(np.abs(xs1[:, r] - xs0[:, r]))) <= 1e-9
    print(f"tol {tol}: {s}")


def test_jacobian_errors():
    c = netlist.parse(DIODE_LADDER).circuit
    for torn, want in [("V1", "branch-current unknown"), ("RL", "connects to ground"), ("RX", "no element")]:
        with pytest.raises(tear.TearError, match=want):
            tear.cut_pairs(c, [torn])
    with pytest.raises(tear.TearError):
        tear.cut_pairs(c, ["RT", "rt"])
    s = netlist.parse("loop\nV1 1 0 1\nV2 1 0 2\nR1 1 2 1k\nR2 2 0 1k\n.end\n").circuit
    o = oracle.options(no_gmin_stepping=1, no_source_stepping=1, no_pseudo_transient=1)
    els, p = s.flatten()
    st_ref, _, _ = oracle.op(s.num_nodes, els, p, opts=o)
    with oracle.Tearing(tear.cut_pairs(s, ["R1"])):
        st, _, _ = oracle.op(s.num_nodes, els, p, opts=o)
    assert st_ref == 1 and st == 1  # ErrSingularMatrix either way


def chain8():
    ri, ci, v = [], [], []
    for i in range(8):
        ri.append(i); ci.append(i); v.append(2.0)
        if i > 0 and i != 4:
            ri += [i, i - 1]; ci += [i - 1, i]; v += [-1.0, -1.0]
    ri += [4, 3]; ci += [3, 4]; v += [-0.5, -0.5]
    return ri, ci, v


def test_chain_torn_in_the_middle():
    t = oracle.Tearing([[3, 4]])
    m = t.matrix(8)
    ri, ci, v = chain8()
    b = np.zeros(8)
    b[0] = 1
    for _ in range(3):
        m.clear()
        m.add_batch(ri, ci, v)
        e, x = m.solve(b)
        assert e == 0
    A = np.zeros((8, 8))
    np.add.at(A, (ri, ci), v)
    assert np.max(np.abs(x - np.linalg.solve(A, b))) <= 1e-15
    s = t.stats()
    assert s["max_pieces"] == 2 and s["fallbacks"] == 0 and s["factors"] == 2 and s["reuses"] == 4, s
    k, po = m.pieces()
    assert k == 2 and list(po) == [0, 0, 0, 0, 1, 1, 1, 1]


def clustered(k, m, rng, complex_=False, coupling=1.0):
    """k chains of m unknowns with dominant diagonals, consecutive chains
    coupled by one symmetric pair each; returns entries and the cut."""
    n = k * m
    ents = {}
    for q in range(k):
        b = q * m
        for i in range(m):
            ents[(b + i, b + i)] = rng.uniform(4, 8)
            if i:
                ents[(b + i, b + i - 1)] = rng.uniform(-1, 1)
                ents[(b + i - 1, b + i)] = rng.uniform(-1, 1)
    cut = []
    for q in range(k - 1):
        a, c = q * m + rng.integers(m), (q + 1) * m + rng.integers(m)
        ents[(a, c)] = coupling * rng.uniform(-1, 1)
        ents[(c, a)] = coupling * rng.uniform(-1, 1)
        cut.append([a, c])
    keys = list(ents)
    ri = np.array([k_[0] for k_ in keys], np.int32)
    ci = np.array([k_[1] for k_ in keys], np.int32)
    v = np.array([ents[k_] for k_ in keys])
    if complex_:
        v = v + 1j * rng.uniform(-0.3, 0.3, len(v))
    return n, ri, ci, v, np.array(cut, np.int32)


@pytest.mark.parametrize("complex_", [False, True])
def test_torn_equals_direct_solve(complex_):
    rng = np.random.default_rng(7)
    for k, m in [(2, 5), (4, 6), (6, 9)]:
        n, ri, ci, v, cut = clustered(k, m, rng, complex_)
        t = oracle.Tearing(cut)
        mat = t.matrix(n, complex_)
        mat.add_batch(ri, ci, v)
        b = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if complex_ else 0)
        e, x = mat.solve(b)
        A = np.zeros((n, n), complex if complex_ else float)
        np.add.at(A, (ri, ci), v)
        want = np.linalg.solve(A, b)
        assert e == 0
        assert np.max(np.abs(x - want)) <= 1e-13 * np.max(np.abs(want))
        s = t.stats()
        assert s["max_pieces"] == k and s["refines"] == 0 and s["fallbacks"] == 0, s


def test_badly_scaled_pieces():
    """Rows scaled over twelve decades: the equilibration keeps the torn
    solve accurate without refinement (componentwise backward error)."""
    rng = np.random.default_rng(3)
    n, ri, ci, v, cut = clustered(3, 8, rng)
    rs = 10.0 ** rng.uniform(-6, 6, n)
    v = v * rs[ri]
    t = oracle.Tearing(cut)
    mat = t.matrix(n)
    mat.add_batch(ri, ci, v)
    b = rng.uniform(-1, 1, n) * rs
    e, x = mat.solve(b)
    assert e == 0
    A = np.zeros((n, n))
    np.add.at(A, (ri, ci), v)
    r = np.abs(A @ x - b) / (np.abs(A) @ np.abs(x) + np.abs(b))
    assert r.max() <= 1e-14
    assert t.stats()["fallbacks"] == 0


def test_singular_piece_falls_back():
    """A piece with an empty equation is singular: the system is solved by
    the untorn LU (solver.Auto: dense here), whose answer and error it returns."""
    rng = np.random.default_rng(5)
    n, ri, ci, v, cut = clustered(3, 6, rng)
    v = v.copy()
    v[ri == 2] = 0.0  # row 2 of the first piece: all zero -> the piece is singular, and so is J
    t = oracle.Tearing(cut)
    mat = t.matrix(n)
    mat.add_batch(ri, ci, v)
    e, _ = mat.solve(np.ones(n))
    assert e == 1 and t.stats()["fallbacks"] == 1
    # a piece singular on its own while J is not: the cut pair's diagonal share
    # is its only entry in that row
    ri2 = np.array([0, 1, 0, 1], np.int32)
    ci2 = np.array([1, 0, 0, 1], np.int32)
    t2 = oracle.Tearing([[0, 1]])
    m2 = t2.matrix(2)
    m2.add_batch(ri2, ci2, [1.0, 1.0, 0.0, 0.0])
    e, x = m2.solve([1.0, 2.0])
    s = t2.stats()
    assert e == 0 and s["fallbacks"] == 1 and np.allclose(x, [2.0, 1.0]), (e, x, s)
    ed, xd = oracle.dense_solve(2, ri2, ci2, [1.0, 1.0, 0.0, 0.0], [1.0, 2.0])
    assert ed == 0 and np.array_equal(x, xd)


def test_nonfinite_is_not_a_fallback():
    t = oracle.Tearing([[3, 4]])
    m = t.matrix(8)
    ri, ci, v = chain8()
    v[0] = math.nan
    m.add_batch(ri, ci, v)
    e, _ = m.solve(np.ones(8))
    assert e == 2 and t.stats()["fallbacks"] == 0 and t.stats()["solves"] == 1
