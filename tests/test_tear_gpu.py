"""JacobianTearing on the GPU (kh_tear_*, kh_tmat_*, kh_tear_batch_*; device
code kronspice_amd/csrc/kh_tear.h) against the oracle's restatement of
kron.JacobianTearing (oracle/tear.hpp, pinned in tests/test_tear_oracle.py).

Bars, all bit for bit:
  * the drop-in seam: every golden netlist's analyses run by the oracle with
    Options.Solver = a JacobianTearing of each single-element cut and the
    all-element cut; the sequence of systems each torn matrix solved (real
    for OP, DC sweeps and transients, complex for AC) is replayed through
    JacobianTearing.Matrix / CMatrix on the GPU — the calls a Go analysis
    makes through the seam — and every solution, status and counter (solves,
    pieces factored and reused, fallbacks, refinement steps, pieces) equals
    the oracle's;
  * the batched form: B systems of one pattern, each with its own values and
    its own piece factors over a sequence of calls (unchanged values, one
    piece changed, moves within the reuse tolerance, singular pieces that
    fall back to the untorn LU — dense below 30 unknowns, sparse above — and
    NaNs), every system equal to an oracle matrix that saw the same calls;
  * a piece over solver.MaxDense: the plan is refused and every solve is
    untorn.
"""
import numpy as np
import pytest

from kronspice_amd import api, tear
from oracle import oracle
from tests import golden_runner as gr
from tests.test_tear_oracle import TornBackend, clustered

pytestmark = pytest.mark.gpu

STAT_KEYS = [("solves", "solves"), ("factors", "factors"), ("reuses", "reuses"), ("stale_reuses", "stale"),
             ("fallbacks", "fallbacks"), ("refinements", "refines"), ("max_pieces", "max_pieces")]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a.view(np.float64).view(np.uint64)


def replay(records, cuts, tol=0.0):
    """The recorded solves through the GPU seam, one torn matrix per recorded
    matrix; returns the GPU tearing's counters."""
    jt = api.JacobianTearing(cuts, tol)
    mats = {}
    for k, r in enumerate(records):
        m = mats.get(r["mat"])
        if m is None:
            m = mats[r["mat"]] = jt.cmatrix(r["n"]) if r["cplx"] else jt.matrix(r["n"])
        m.clear()
        m.add_batch(r["ri"], r["ci"], r["val"])
        try:
            x, st = m.solve(r["b"]), 0
        except api.KhError as e:
            x, st = None, e.status
        assert st == r["status"], (k, st, r["status"])
        if st == 0:
            assert np.array_equal(bits(x), bits(r["x"])), (k, np.max(np.abs(x - r["x"])))
    return jt.stats()


def same_stats(g, o):
    for kg, ko in STAT_KEYS:
        assert g[kg] == o[ko], (kg, g, o)


@pytest.mark.parametrize("name", gr.names())
def test_seam_replays_goldens_under_every_cut(name):
    c = gr.load_deck(name).circuit
    total = 0
    for cut in tear.jacobian_cuts(c):
        b = TornBackend(gr.load_deck(name).circuit, cut)
        with oracle.TearRecording() as rec:
            msgs = gr.check(name, b)
        assert not msgs, msgs[:3]
        g = replay(rec.records, b.t.cuts)
        same_stats(g, b.t.stats())
        total += len(rec.records)
    print(f"{name}: {len(tear.jacobian_cuts(c))} cuts, {total} torn solves replayed")


def test_seam_reuse_tolerance_replay():
    """The diode ladder torn from its rectifier with ReuseTol 0.01: stale
    reuses, refinement steps and the second torn pass all agree."""
    from kronspice_amd import netlist
    from tests.test_tear_oracle import DIODE_LADDER, tran_of
    c = netlist.parse(DIODE_LADDER).circuit
    t = oracle.Tearing(tear.cut_pairs(c, ["RT"]), 0.01)
    with oracle.TearRecording() as rec:
        st, *_ = tran_of(c, t)
    assert st == 0
    o = t.stats()
    assert o["stale"] > 0 and o["refines"] > 0
    same_stats(replay(rec.records, t.cuts, 0.01), o)


def run_batch(n, ri, ci, B, cuts, calls, tol, complex_):
    """calls: list of (vals [nslots, B], b [n, B]) numpy arrays. Each call goes
    through the GPU batch and, system by system, through B oracle matrices."""
    torch = api._torch()
    jt = api.JacobianTearing(cuts, tol)
    tb = jt.batch(n, ri, ci, B, complex_)
    ot = oracle.Tearing(cuts, tol)
    om = [ot.matrix(n, complex_) for _ in range(B)]
    dt = torch.complex128 if complex_ else torch.float64
    summary = []
    for k, (vals, b) in enumerate(calls):
        dv = torch.from_numpy(np.ascontiguousarray(vals)).to("cuda", dt)
        db = torch.from_numpy(np.ascontiguousarray(b)).to("cuda", dt)
        dx = torch.zeros((n, B), dtype=dt, device="cuda")
        ds = torch.full((B,), -1, dtype=torch.int32, device="cuda")
        tb.solve(dv, db, dx, ds)
        torch.cuda.synchronize()
        x, st = dx.cpu().numpy(), ds.cpu().numpy()
        for c in range(B):
            om[c].clear()
            om[c].add_batch(ri, ci, vals[:, c])
            e, xo = om[c].solve(b[:, c])
            assert st[c] == e, (k, c, st[c], e)
            if e == 0:
                assert np.array_equal(bits(x[:, c]), bits(xo)), (k, c)
        summary.append(int((st == 0).sum()))
    g, o = jt.stats(), ot.stats()
    same_stats(g, o)
    k, po = tb.pieces()
    return g, summary, k


def sequence(n, ri, ci, B, rng, complex_, m):
    base = np.abs(rng.uniform(0.5, 1.5, (len(ri), B)))
    v0 = np.empty((len(ri), B), complex if complex_ else float)
    for s in range(len(ri)):
        v0[s] = (rng.uniform(4, 8, B) if ri[s] == ci[s] else rng.uniform(-1, 1, B)) * base[s]
        if complex_:
            v0[s] = v0[s] + 1j * rng.uniform(-0.3, 0.3, B)
    b = rng.uniform(-1, 1, (n, B)) + (1j * rng.uniform(-1, 1, (n, B)) if complex_ else 0)
    calls = [(v0, b), (v0, b)]
    v1 = v0.copy()
    v1[(ri == ci) & (ri // m == 1)] *= 1.25  # one piece changes in every system
    calls.append((v1, b))
    calls.append((v1 * (1 + 1e-6), b))  # every value moves a little
    v2 = v1.copy()
    v2[ri == 0, 1::2] = 0  # row 0 empty in odd systems: a singular piece, and a singular system
    calls.append((v2, b))
    v3 = v1.copy()
    v3[(ri == m + 1) & (ci == m + 1), 2::3] = 0  # a zero diagonal in every third system's second piece
    v3[(ri == m + 1) & (ci == m), 2::3] *= 1e-9
    calls.append((v3, b))
    v4 = v1.copy()
    v4[len(ri) - 1, ::4] = np.nan
    calls.append((v4, b))
    calls.append((v1, b))
    return calls


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("k,m,tol", [(4, 6, 0.0), (4, 6, 0.01), (3, 40, 0.0), (6, 9, 0.0)])
def test_batch_matches_oracle(k, m, tol, complex_):
    rng = np.random.default_rng(k * 100 + m)
    n, ri, ci, v, cut = clustered(k, m, rng, complex_)
    B = 96
    calls = sequence(n, ri, ci, B, rng, complex_, m)
    g, ok, pieces = run_batch(n, ri, ci, B, cut, calls, tol, complex_)
    print(f"n={n} pieces={pieces} tol={tol} complex={complex_}: solved per call {ok}; {g}")
    assert pieces == k and g["fallbacks"] > 0 and g["reuses"] > 0
    if tol > 0:
        assert g["stale_reuses"] > 0


def test_refused_plan_is_untorn():
    """One piece of 4,097 unknowns (a chain, torn nowhere): larger than
    solver.MaxDense, so the plan is refused and each solve is the untorn
    sparse LU's."""
    n = 4097
    ri = np.concatenate([np.arange(n), np.arange(1, n), np.arange(n - 1)]).astype(np.int32)
    ci = np.concatenate([np.arange(n), np.arange(n - 1), np.arange(1, n)]).astype(np.int32)
    v = np.concatenate([np.full(n, 4.0), np.full(2 * (n - 1), -1.0)])
    b = np.linspace(-1, 1, n)
    jt = api.JacobianTearing(np.zeros((0, 2), np.int32))
    m = jt.matrix(n)
    ot = oracle.Tearing(np.zeros((0, 2), np.int32))
    om = ot.matrix(n)
    for _ in range(2):
        m.clear(); m.add_batch(ri, ci, v)
        om.clear(); om.add_batch(ri, ci, v)
        x = m.solve(b)
        e, xo = om.solve(b)
        assert e == 0 and np.array_equal(bits(x), bits(xo))
    same_stats(jt.stats(), ot.stats())
    assert jt.stats()["fallbacks"] == 2 and jt.stats()["max_pieces"] == 0
    assert m.pieces()[0] == -1
