"""The DC fallback chain: gmin stepping, source stepping and pseudo-transient
continuation (analysis/newton.go:95-311), and the LTE step proposal
(device/lte.go:50-85).

Oracle pins (CPU):
  * analysis/newton_test.go:33-44 `hardDiodes` deck: the operating point and
    its ledger (:46-62), the typed no-convergence error (:64-72), and each
    stepping method run alone reaching plain Newton's point within
    1e-3·|x| + 1e-5 (:74-104).
  * known answers for each continuation schedule on a circuit where every
    Newton solve converges: gmin 1e-2, 1e-3, ... down past Options.Gmin then
    0 (newton.go:140-158: the factor starts at 10 and min(f·√f, 10) keeps it
    at 10); source factors 0, 0.1, 0.25, 0.475, 0.8125, 1 (newton.go:190-207:
    the step starts at 0.1 and grows ×1.5 up to 0.5); pseudo-transient steps
    1 µs·2^k until the step reaches 1e6 s, then a plain solve (newton.go:
    268-309). Each known answer catches a changed constant, which planted
    mistakes in the oracle demonstrate.
  * the retry rules on the hard deck with ITL1 = 2..4: a failed gmin step
    retries between the last good and the failed conductance, a failed
    source step halves the increment, a failed pseudo-transient step divides
    h by 4.
  * lteStep's closed form: a charge history that is a polynomial of degree
    k+1 has (k+1)-th divided difference equal to its leading coefficient a,
    so the proposal is trtol·(reltol·max|i| + abstol)/a for backward Euler
    and √(2·trtol·(reltol·max|i| + abstol)/a) for the trapezoidal rule,
    whatever the candidate step.

GPU parity: a batch of 64 variants of the hard deck with ITL1 low enough that
plain Newton fails, each fallback method enabled alone: status, Newton
iteration counts and polished operating points against the oracle.
"""
import math

import numpy as np
import pytest

from kronspice_amd import netlist
from oracle import oracle

HARD_DIODES = """hard
V1 in 0 100
R1 in a 10
D1 a b dd
D2 b c dd
D3 c 0 dd
R2 a z 100
DZ 0 z dz
.model dd D(IS=1e-15 N=1 RS=0.1)
.model dz D(IS=1e-14 BV=5.6 IBV=5m RS=1)
.op
"""

EASY_DIODE = """easy
V1 a 0 5
R1 a b 1k
D1 b 0 dm
.model dm D(IS=1e-14)
.op
"""

METHOD_FLAGS = {
    "gmin": dict(no_source_stepping=1, no_pseudo_transient=1),
    "source": dict(no_gmin_stepping=1, no_pseudo_transient=1),
    "ptran": dict(no_gmin_stepping=1, no_source_stepping=1),
}


def deck(text):
    c = netlist.parse(text).circuit
    els, p = c.flatten()
    return c, els, p


# ----------------------------------------------------------------------------- oracle pins
def test_hard_diodes_operating_point():
    """analysis/newton_test.go:46-62: D1 carries over 1 A, the zener clamps
    between 5 and 5.6 V, the ledger closes to 1e-12."""
    c, els, p = deck(HARD_DIODES)
    st, x, info = oracle.op(c.num_nodes, els, p)
    assert st == 0
    assert info["ledger_rel"] <= 1e-12
    assert info["current"][c.element_index("D1")] >= 1
    vz = x[c.node("z") - 1]
    assert 5 <= vz <= 5.6, vz


def test_hard_diodes_no_convergence_is_typed():
    """analysis/newton_test.go:64-72."""
    c, els, p = deck(HARD_DIODES)
    o = oracle.options(itl1=2, no_gmin_stepping=1, no_source_stepping=1, no_pseudo_transient=1)
    st, _, _ = oracle.op(c.num_nodes, els, p, opts=o)
    assert st == 4


@pytest.mark.parametrize("method", ["gmin", "source", "ptran"])
def test_stepping_agrees_with_newton(method):
    """analysis/newton_test.go:74-104: each method alone, from a fresh
    circuit, reaches plain Newton's operating point."""
    c, els, p = deck(HARD_DIODES)
    st0, x0, _ = oracle.dc_method(c.num_nodes, els, p, "newton")
    assert st0 == 0
    st, x, it = oracle.dc_method(c.num_nodes, els, p, method)
    assert st == 0, method
    assert it > 0
    assert np.all(np.abs(x - x0) <= 1e-3 * np.abs(x0) + 1e-5), (method, x, x0)


def _schedule(method, text=EASY_DIODE):
    c, els, p = deck(text)
    st, x, it, tr = oracle.dc_method(c.num_nodes, els, p, method, trace=True)
    return st, tr


def test_gmin_schedule_known_answer():
    st, tr = _schedule("gmin")
    assert st == 0
    assert (tr[:-1, 0] == 1).all() and (tr[:, 2] == 0).all()
    assert tr[-1, 0] == 0  # g = 0: the last solve has no diagonal conductance
    g = tr[:, 1]
    want = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12, 0.0]
    assert len(g) == len(want), g
    assert np.allclose(g, want, rtol=1e-12, atol=0), g


def test_source_schedule_known_answer():
    st, tr = _schedule("source")
    assert st == 0
    assert (tr[:, 0] == 2).all() and (tr[:, 2] == 0).all()
    assert len(tr) == 6, tr[:, 1]
    assert np.allclose(tr[:, 1], [0.0, 0.1, 0.25, 0.475, 0.8125, 1.0], rtol=1e-15, atol=0), tr[:, 1]
    assert tr[-1, 1] == 1.0


def test_pseudo_transient_schedule_known_answer():
    st, tr = _schedule("ptran")
    assert st == 0
    h = tr[:-1, 1]
    assert (tr[:-1, 0] == 3).all() and (tr[:, 2] == 0).all()
    assert len(h) == 41  # 1e-6 * 2**40 is the first step >= 1e6 s
    assert np.array_equal(h, 1e-6 * 2.0 ** np.arange(41))
    assert tr[-1, 0] == 0  # the confirming plain Newton solve


@pytest.mark.parametrize("which,value,method", [("gmin_cap", 100.0, "gmin"), ("source_growth", 2.0, "source"),
                                                ("pt_growth", 3.0, "ptran")])
def test_planted_stepping_mistakes_are_caught(which, value, method):
    """A changed stepping constant in the oracle breaks the known answer of
    its schedule (the mistake is planted in the oracle, on the CPU)."""
    checks = {"gmin": test_gmin_schedule_known_answer, "source": test_source_schedule_known_answer,
              "ptran": test_pseudo_transient_schedule_known_answer}
    oracle.plant_stepping(which, value)
    try:
        with pytest.raises(AssertionError):
            checks[method]()
    finally:
        oracle.plant_stepping(which, 0.0)
    checks[method]()


@pytest.mark.parametrize("itl1", [2, 3, 4])
def test_stepping_retry_rules(itl1):
    c, els, p = deck(HARD_DIODES)
    seen_fail = 0
    for method in ("gmin", "source", "ptran"):
        o = oracle.options(itl1=itl1)
        st, x, it, tr = oracle.dc_method(c.num_nodes, els, p, method, opts=o, trace=True)
        kind = {"gmin": 1, "source": 2, "ptran": 3}[method]
        rows = tr[tr[:, 0] == kind]
        good = None
        for k in range(len(rows) - 1):
            par, status = rows[k, 1], rows[k, 2]
            nxt = rows[k + 1, 1]
            if status == 0:
                good = par
                continue
            seen_fail += 1
            if method == "gmin" and good is not None:
                # newton.go:166-173: retry between the last good g and the failed one
                lo, hi = (par, good) if par < good else (good, par)
                assert lo <= nxt <= hi or (par == 0 and nxt > 0), (k, par, good, nxt)
            elif method == "source":
                # newton.go:208-209: the increment from the last good factor
                # halves (a factor capped at 1 hid the full increment)
                assert good < nxt <= par, (k, par, good, nxt)
                if par < 1:
                    assert nxt - good == pytest.approx((par - good) / 2, rel=1e-12), (k, par, good, nxt)
            elif method == "ptran":
                assert nxt == par / 4, (k, par, nxt)  # newton.go:284
    assert seen_fail > 0


@pytest.mark.parametrize("method,a", [(1, 2.0e3), (2, 5.0e5)])
@pytest.mark.parametrize("h", [1e-5, 4e-5])
def test_lte_step_closed_form(method, a, h):
    k = method
    t0 = 1e-3
    ts = [t0 - h, t0 - h - 1.5e-5, t0 - h - 3.5e-5]
    q = lambda t: a * t ** (k + 1) + 3e-7 * t + 2e-11  # noqa: E731
    i_n, i_prev = 4000.0, 2500.0
    reltol, abstol, trtol, chgtol = 1e-3, 1e-12, 7.0, 1e-14
    hn = oracle.lte_step(k, t0, h, ts, [q(t) for t in ts], 3, q(t0), i_n, i_prev, abstol, reltol, chgtol, trtol)
    qmax = max(abs(q(t0)), abs(q(ts[0])))
    cur = reltol * max(abs(i_n), abs(i_prev)) + abstol
    assert reltol * max(qmax, chgtol) / h < cur  # the current term decides the tolerance
    want = trtol * cur / a if k == 1 else math.sqrt(2 * trtol * cur / a)
    assert hn == pytest.approx(want, rel=1e-6), (hn, want)
    # too little history: no proposal
    assert oracle.lte_step(k, t0, h, ts, [q(t) for t in ts], k, q(t0), i_n, i_prev) == math.inf


# ----------------------------------------------------------------------------- GPU parity
def hard_batch(B, seed=11):
    """B variants of the hard deck: V1 U(50, 150) V, R1 U(5, 20) ohm, R2 U(50, 200) ohm."""
    rng = np.random.default_rng(seed)
    c = netlist.parse(HARD_DIODES).circuit
    kv, k1, k2 = c.element_index("V1"), c.element_index("R1"), c.element_index("R2")
    cols = []
    els = None
    for _ in range(B):
        c.elements[kv].value = float(rng.uniform(50, 150))
        c.elements[k1].value = float(rng.uniform(5, 20))
        c.elements[k2].value = float(rng.uniform(50, 200))
        e, p = c.flatten()
        if els is None:
            els = e
        cols.append(p)
    return c, els, np.stack(cols, axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 2])
@pytest.mark.parametrize("method,itl1", [("gmin", 4), ("gmin", 6), ("source", 4), ("source", 8), ("ptran", 2),
                                         ("ptran", 5)])
def test_gpu_fallback_chain_matches_oracle(method, itl1, backend):
    from kronspice_amd import api
    c, els, P = hard_batch(64)
    kw = dict(itl1=itl1, backend=backend, **METHOD_FLAGS[method])
    t = api.Topology(c.num_nodes, els, P[:, 0])
    o = api.Options(**kw)
    r = api.op(t, P, opts=o, polish=True)
    x, st, it = r.x.cpu().numpy(), r.status.cpu().numpy(), r.iters.cpu().numpy()
    oo = oracle.options(**kw)
    plain = oracle.options(itl1=itl1, backend=backend, no_gmin_stepping=1, no_source_stepping=1,
                           no_pseudo_transient=1)
    fell_back = 0
    for b in range(P.shape[1]):
        e0, x0, it0 = oracle.op_polished(c.num_nodes, els, np.ascontiguousarray(P[:, b]), opts=oo)
        assert st[b] == e0, (b, st[b], e0)
        assert it[b] == it0, (b, it[b], it0)
        if e0 == 0:
            rel = np.max(np.abs(x[:, b] - x0)) / np.max(np.abs(x0))
            assert rel <= 1e-9, (b, rel)
        ep, _, _ = oracle.op(c.num_nodes, els, np.ascontiguousarray(P[:, b]), opts=plain)
        fell_back += ep != 0
    # plain Newton fails on a good share of the batch at this ITL1 (on all of it
    # at ITL1 = 2): there the fallback method did the work
    assert fell_back >= P.shape[1] // 4, fell_back
    assert (st == 0).mean() >= 0.8  # the oracle agrees on every status (checked above)
