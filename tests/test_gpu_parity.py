"""Parity of the GPU solve path (libkronhip.so, through its C ABI) with the
FP64 oracle and with the reference's own fixtures.

Bars: bit-exact pivot rows, L/U fill and solutions wherever the computation
is +,-,*,/ only (linear circuits, the LU itself); for circuits whose device
models call exp/log/sin (the GPU's math library differs from glibc in the
last ulp) the same status, the same Newton iteration and time-step counts on
at least 99 % of circuits, and node voltages / branch currents within 1e-9
relative to the oracle's (the reference's torn-vs-untorn bound,
analysis/backend_test.go:59-117).
"""
import os

import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle
from tests import golden_runner as gr

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    return api.to_device(a, dtype)


def host(t):
    return t.cpu().numpy()


def relerr(a, b):
    """max |a - b| / max |b| per column (circuit)."""
    den = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return np.max(np.abs(a - b), axis=0) / den


# ---------------------------------------------------------------------------- solver seam
def test_seam_zero_diagonal_exact():
    s = api.Sparse(2)
    s.add(0, 0, 1e-3); s.add(0, 1, 1); s.add(1, 0, 1)
    x = s.solve([0, 5])
    assert abs(x[0] - 5) <= 1e-12 and abs(x[1] + 5e-3) <= 1e-15


@pytest.mark.parametrize("case", ["empty", "zero column", "dependent rows", "nan"])
def test_seam_error_classes(case):
    s = api.Sparse(3)
    want = api.SINGULAR
    if case == "zero column":
        s.add(0, 0, 1); s.add(1, 0, 1); s.add(2, 2, 1)
    elif case == "dependent rows":
        for j in range(3):
            s.add(0, j, 1); s.add(1, j, 2)
        s.add(2, 2, 1)
    elif case == "nan":
        s.add(0, 0, float("nan")); s.add(1, 1, 1); s.add(2, 2, 1)
        want = api.NONFINITE
    with pytest.raises(api.KhError) as e:
        s.solve(np.zeros(3))
    assert e.value.status == want


def test_seam_ground_and_growth():
    s = api.Sparse(1)
    s.add(-1, 0, 5); s.add(0, -1, 5); s.add(0, 0, 2)
    assert s.solve([4])[0] == 2
    s = api.Sparse(2)
    s.add(0, 0, 2); s.add(1, 1, 3)
    x = s.solve([2, 3])
    assert x[0] == 1 and x[1] == 1
    s.add(0, 1, 1); s.add(1, 0, -1)
    x = s.solve([2, 3])
    o = oracle.Sparse(2)
    for i, j, v in [(0, 0, 2), (1, 1, 3), (0, 1, 1), (1, 0, -1)]:
        o.add(i, j, v)
    assert np.array_equal(x, o.solve([2, 3])[1])


def grid(s, w, h, nv, rng):
    nn = w * h
    def g(a, b, v):
        s.add(a, a, v); s.add(b, b, v); s.add(a, b, -v); s.add(b, a, -v)
    for y in range(h):
        for x in range(w):
            i = y * w + x
            if x + 1 < w:
                g(i, i + 1, 1 / (10 + 990 * rng.random()))
            if y + 1 < h:
                g(i, i + w, 1 / (10 + 990 * rng.random()))
            s.add(i, i, 1e-9)
    b = np.zeros(nn + nv)
    for k in range(nv):
        node, br = (k * 7919) % nn, nn + k
        s.add(node, br, 1); s.add(br, node, 1)
        b[br] = 1 + rng.random()
    return b


def test_seam_fill_40x40_matches_reference_counts():
    # doc/M6.md:77-78: nnz A 7,848, L 19,880, U 21,493 for gridMNA(40, 40, 4).
    s = api.Sparse(1604)
    b = grid(s, 40, 40, 4, np.random.default_rng(5))
    x = s.solve(b)
    assert s.nnz() == (7848, 19880, 21493)
    o = oracle.Sparse(1604)
    grid(o, 40, 40, 4, np.random.default_rng(5))
    e, xo = o.solve(b)
    q, prow = s.pivots()
    info = o.info()
    assert np.array_equal(q, info["q"]) and np.array_equal(prow, info["prow"])
    assert np.array_equal(x, xo)


def test_seam_random_unsymmetric_bitexact():
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = 1 + int(rng.integers(60))
        s, o = api.Sparse(n), oracle.Sparse(n)
        ents = [(i, (i + int(rng.integers(n))) % n, 1.0) for i in range(n)]
        ents += [(int(rng.integers(n)), int(rng.integers(n)), rng.normal() * 0.1) for _ in range(3 * n)]
        for i, j, v in ents:
            s.add(i, j, v); o.add(i, j, v)
        b = rng.normal(size=n)
        e, xo = o.solve(b)
        if e != 0:
            with pytest.raises(api.KhError):
                s.solve(b)
            continue
        x = s.solve(b)
        assert np.array_equal(x, xo), trial
        assert np.array_equal(s.pivots()[1], o.info()["prow"]), trial


# ---------------------------------------------------------------------------- batched LU
def test_lu_batch_bitexact_pivots_and_solution():
    rng = np.random.default_rng(11)
    n, B = 66, 512
    # A ladder-shaped MNA pattern with a branch row (zero diagonal): node 0
    # driven by a source, series and shunt conductances with random values.
    ri, ci = [], []
    def ent(i, j):
        ri.append(i); ci.append(j)
    nn = n - 1
    for i in range(nn):
        ent(i, i)
        if i + 1 < nn:
            ent(i, i + 1); ent(i + 1, i)
    ent(0, nn); ent(nn, 0)
    nnz = len(ri)
    vals = np.zeros((nnz, B))
    g = np.exp(np.log(1e-5) + rng.random((2 * nn, B)) * np.log(1e3))
    k = 0
    for idx, (i, j) in enumerate(zip(ri, ci)):
        if i == j:
            v = g[i] + (g[nn + i - 1] if i > 0 else 0) + (g[nn + i] if i + 1 < nn else 0)
        elif i == nn or j == nn:
            v = np.ones(B)
        else:
            v = -g[nn + min(i, j)]
        vals[idx] = v
    rhs = np.zeros((n, B)); rhs[nn] = 10.0
    lu = api.LUBatch(n, ri, ci, B)
    dv, db = dev(vals), dev(rhs)
    import torch
    x = torch.zeros((n, B), dtype=torch.float64, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    prow = torch.zeros((n, B), dtype=torch.int32, device="cuda")
    lu.solve(dv, db, x, st, prow)
    torch.cuda.synchronize()
    xg, sg, pg = host(x), host(st), host(prow)
    assert (sg == 0).all()
    for c in range(0, B, 7):
        o = oracle.Sparse(n)
        for idx, (i, j) in enumerate(zip(ri, ci)):
            o.add(i, j, vals[idx, c])
        e, xo = o.solve(rhs[:, c])
        assert e == 0
        assert np.array_equal(xg[:, c], xo), c
        assert np.array_equal(pg[:, c], o.info()["prow"]), c


# ---------------------------------------------------------------------------- analyses
@pytest.mark.parametrize("name", [n for n in gr.names() if "tran" not in n and "_ac" not in n])
def test_op_goldens_bitexact_with_oracle(name):
    deck = gr.load_deck(name)
    c = deck.circuit
    kinds = [a.kind for a in deck.analyses]
    points = [None]
    if "dc" in kinds:
        a = [a for a in deck.analyses if a.kind == "dc"][0].args
        k = c.element_index(a["source"])
        points = list(np.arange(a["start"], a["stop"] + a["step"] / 2, a["step"]))[::4]
    for v in points:
        if v is not None:
            c.elements[k].value = float(v)
        els, p = c.flatten()
        t = api.Topology(c.num_nodes, els, p)
        B = 64
        r = api.op(t, np.repeat(p[:, None], B, axis=1))
        x, st, it = host(r.x), host(r.status), host(r.iters)
        e0, x0, info = oracle.op(c.num_nodes, els, p)
        assert (st == e0).all(), (v, st[0], e0)
        assert (it == info["iters"]).all(), (v, it[0], info["iters"])
        rel = relerr(x, np.repeat(x0[:, None], B, axis=1))
        linear = all(e.kind not in (6, 7, 8) for e in c.elements)
        if linear:
            assert (x == x0[:, None]).all(), v
        else:
            assert rel.max() <= 1e-9, (v, rel.max())


@pytest.mark.parametrize("name", [n for n in gr.names() if "tran" in n or n in ("rc_step", "rl_step", "rlc_step", "rectifier_rc")])
def test_tran_goldens_same_steps_as_oracle(name):
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "tran"][0].args
    els, p = c.flatten(a["step"], a["stop"])
    spec = api.TranSpec(step=a["step"], stop=a["stop"], start=a["start"], max_step=a["max_step"],
                        uic=1 if a["uic"] else 0)
    t = api.Topology(c.num_nodes, els, p)
    r = api.tran(t, p[:, None], spec)
    ospec = oracle.KhTranSpec(step=a["step"], stop=a["stop"], start=a["start"], max_step=a["max_step"],
                              uic=1 if a["uic"] else 0)
    e0, x0, info = oracle.tran(c.num_nodes, els, p, ospec)
    assert int(host(r.status)[0]) == e0
    assert int(host(r.accepted)[0]) == info["accepted"]
    assert int(host(r.rejected)[0]) == info["rejected"]
    x = host(r.x)[:, 0]
    assert np.max(np.abs(x - x0)) <= 1e-9 * np.max(np.abs(x0))


def test_ladder_batch_op_bitexact():
    b = W.resistor_ladders(1024)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.op(t, b.params)
    xo, so, io = oracle.batch_op(b.circuit.num_nodes, b.elements, b.params, b.n, threads=8)
    assert (host(r.status) == 0).all() and (so == 0).all()
    assert np.array_equal(host(r.x), xo)
    assert r.stats["patterns"] == 1


def test_rectifier_batch_tran_matches_oracle():
    b = W.rectifier_rc_ladders(256)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.tran(t, b.params, spec, lane_spec=b.lane_spec)
    ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    xo, so, ao, ro, lo = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params, b.n, ospec, threads=8,
                                           lane_spec=b.lane_spec)
    st, acc, rej = host(r.status), host(r.accepted), host(r.rejected)
    assert (st == so).all()
    same = (acc == ao) & (rej == ro)
    assert same.mean() >= 0.99, same.mean()
    rel = relerr(host(r.x), xo)
    assert rel[same].max() <= 1e-9, rel[same].max()
    assert host(r.ledger).max() <= 1e-9


@pytest.mark.parametrize("name", ["diode_fwd", "mos1_op", "mos1_cmos_vtc", "mos1_cs_ac", "bjt_ce_op"])
def test_nodeset_warm_start_matches_oracle(name):
    """Options.NodeSet (analysis/newton.go:221-243): Newton starts from a
    per-circuit guess (node voltages only; branch currents and internal nodes
    start at 0) with InitJct off and device limiting seeded from it. Guesses
    are the converged node voltages scaled by 0.9 .. 1.1; each circuit must
    reach the oracle's status, iteration count and solution from the same
    guess."""
    c = gr.load_deck(name).circuit
    els, p = c.flatten()
    e0, x0, info0 = oracle.op(c.num_nodes, els, p)
    assert e0 == 0
    B = 16
    scale = np.linspace(0.9, 1.1, B)
    scale[B // 2] = 1.0
    guess = x0[: c.num_nodes - 1, None] * scale[None, :]
    t = api.Topology(c.num_nodes, els, p)
    r = api.op(t, np.repeat(p[:, None], B, axis=1), nodeset=np.ascontiguousarray(guess))
    x, st, it = host(r.x), host(r.status), host(r.iters)
    for k in range(B):
        e, xo, info = oracle.op(c.num_nodes, els, p, guess=np.ascontiguousarray(guess[:, k]))
        assert st[k] == e, (k, st[k], e)
        assert it[k] == info["iters"], (k, it[k], info["iters"])
        assert np.max(np.abs(x[:, k] - xo)) <= 1e-9 * np.max(np.abs(xo)), k


def test_nodeset_latch_states_match_oracle():
    """A bistable latch: the guess decides which operating point Newton
    reaches (cold start: the metastable point; Newton's basins are not
    monotonic in the guess). Guesses sweep q over 0..5 V with qb = 5 - q;
    each circuit must end where the oracle does from the same guess, with the
    same iteration count, and both stable states must be reached."""
    c = W.cmos_latch()
    els, p = c.flatten()
    vdd, qb, q = c.node("vdd") - 1, c.node("qb") - 1, c.node("q") - 1
    B = 64
    guess = np.zeros((c.num_nodes - 1, B))
    guess[vdd] = 5.0
    hi = np.linspace(3.0, 5.0, B)
    guess[q] = np.where(np.arange(B) % 2 == 0, hi, 5.0 - hi)
    guess[qb] = 5.0 - guess[q]
    t = api.Topology(c.num_nodes, els, p)
    r = api.op(t, np.repeat(p[:, None], B, axis=1), nodeset=guess)
    x, st, it = host(r.x), host(r.status), host(r.iters)
    for k in range(B):
        e, xo, info = oracle.op(c.num_nodes, els, p, guess=np.ascontiguousarray(guess[:, k]))
        assert st[k] == e == 0 and it[k] == info["iters"], (k, st[k], e, it[k], info["iters"])
        assert np.max(np.abs(x[:, k] - xo)) <= 1e-9 * np.max(np.abs(xo)), k
    assert (x[q] > 4.9).any() and (x[qb] > 4.9).any()  # both stable states are reached


@pytest.mark.parametrize("name", ["diode_clipper_dc", "mos1_idvd_dc", "mos1_cmos_vtc", "zener_dc", "bjt_output_dc",
                                  "bjt_pnp_dc", "mos2_cmos_vtc", "mos2_idvd_dc", "mos2_subvt_dc", "mos3_cmos_vtc",
                                  "mos3_idvd_dc", "mos3_subvt_dc"])
def test_dc_sweep_matches_oracle(name):
    """analysis.DC (analysis/dc.go:51-130) through kh_batch_dc: every point
    warm-started from the previous one on the same circuit state; the same
    points solved, the same Newton iterations at every point and solutions
    within 1e-9 of the oracle's sweep. mos1_idvd_dc also sweeps VGS as the
    outer loop (two sweeps)."""
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "dc"][0].args
    els, p = c.flatten()
    k = c.element_index(a["source"])
    sweeps = [(k, a["start"], a["stop"], a["step"])]
    if name == "mos1_idvd_dc":
        sweeps = [(k, a["start"], a["stop"], 0.5), (c.element_index("vgs"), 1.0, 3.0, 1.0)]
    e0, xo, ito, doneo, mro = oracle.dc(c.num_nodes, els, p, sweeps)
    assert e0 == 0
    B = 8
    t = api.Topology(c.num_nodes, els, p)
    r = api.dc(t, np.repeat(p[:, None], B, axis=1), sweeps)
    x, st, it, done = host(r.x), host(r.status), host(r.iters), host(r.points)
    assert (st == 0).all() and (done == doneo).all()
    for b in range(B):
        assert np.array_equal(it[:, b], ito), b
        assert np.max(np.abs(x[:, :, b] - xo)) <= 1e-9 * np.max(np.abs(xo)), b
    assert host(r.ledger).max() <= 1e-9


AC_DECKS = ["rc_lowpass_ac", "rlc_bandpass_ac", "mos1_cs_ac", "bjt_ce_ac", "mos3_cs_ac", "param_rc_ac", "mutual_ac"]
# Decks whose operating point involves no transcendental function of a value
# that differs between the device library and glibc: bit for bit. The BJT's
# exponentials and the level-3 MOSFET's exponentials and powers may differ in
# the last bit: within 1e-9.
AC_BITEXACT = {"rc_lowpass_ac", "rlc_bandpass_ac", "mos1_cs_ac", "param_rc_ac", "mutual_ac"}


@pytest.mark.parametrize("name", AC_DECKS)
def test_ac_goldens_match_oracle(name):
    """analysis.AC (analysis/ac.go:76-193) through kh_batch_ac: operating point,
    polish, then per frequency the LoadAC stamps, the complex GP solve (CSparse:
    cmplx.Abs pivoting, Go's complex division) and the complex-Tellegen ledger.
    Every frequency's solution matches the oracle's bit for bit (the BJT and
    level-3 decks: within 1e-9, their exponentials being the device
    library's)."""
    from kronspice_amd import netlist
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "ac"][0].args
    freqs = np.array(netlist.ac_frequencies(a["variation"], a["points"], a["fstart"], a["fstop"]))
    els, p = c.flatten()
    e0, xo, mro = oracle.ac(c.num_nodes, els, p, freqs)
    assert e0 == 0
    B = 16
    t = api.Topology(c.num_nodes, els, p)
    r = api.ac(t, np.repeat(p[:, None], B, axis=1), freqs)
    x, st, pts = host(r.x), host(r.status), host(r.points)
    assert (st == 0).all() and (pts == len(freqs)).all()
    for b in range(B):
        if name in AC_BITEXACT:
            assert np.array_equal(x[:, :, b], xo), b
        else:
            assert np.max(np.abs(x[:, :, b] - xo)) <= 1e-9 * np.max(np.abs(xo)), b
    assert host(r.ledger).max() <= 1e-9


@pytest.mark.parametrize("backend", [0, 2])
def test_ac_bjt_sweep_from_0hz_grows_pattern(backend):
    """A linear sweep that starts at 0 Hz (ACSpec lin, analysis/ac.go:32-73):
    the BJT's base-collector admittance jwC_bx (RB > 0, XCJC < 1) is zero at
    0 Hz, so LoadAC skips it there (device/bjt.go:721-734) and the reference's
    CSparse adds those entries at the first nonzero frequency, re-ordering the
    grown pattern (solver/sparse.go:114-143). Every frequency, forced sparse
    or under Auto, within 1e-9 of the oracle, which grows its pattern the same
    way."""
    deck = gr.load_deck("bjt_ce_ac")
    c = deck.circuit
    els, p = c.flatten()
    freqs = np.linspace(0.0, 2e8, 9)
    o = oracle.options(backend=backend)
    e0, xo, _ = oracle.ac(c.num_nodes, els, p, freqs, opts=o)
    assert e0 == 0
    t = api.Topology(c.num_nodes, els, p)
    B = 8
    r = api.ac(t, np.repeat(p[:, None], B, axis=1), freqs, opts=api.Options(backend=backend))
    x, st, pts = host(r.x), host(r.status), host(r.points)
    assert (st == 0).all() and (pts == len(freqs)).all()
    for b in range(B):
        assert np.max(np.abs(x[:, :, b] - xo)) <= 1e-9 * np.max(np.abs(xo)), b


def test_ac_rc_lowpass_batch_closed_form():
    """Many RC low-pass circuits with random R and C: v(out)/v(in) must be the
    textbook 1/(1 + j w R C) at every frequency (to 1e-12), and every circuit
    equal to the oracle's sweep of the same values."""
    from kronspice_amd import circuit as kc
    B, nf = 256, 64
    rng = np.random.default_rng(5)
    R = np.exp(rng.uniform(np.log(10.0), np.log(1e5), B))
    Cv = np.exp(rng.uniform(np.log(1e-12), np.log(1e-6), B))
    base = (kc.Builder("rc").vsource("V1", "in", "0", 0.0, ac=1.0).resistor("R1", "in", "out", 1e3)
            .capacitor("C1", "out", "0", 1e-9).build())
    els, p0 = base.flatten()
    params = np.repeat(p0[:, None], B, axis=1)
    kr, kcap = base.element_index("R1"), base.element_index("C1")
    params[els[kr]["poff"]] = R
    params[els[kr]["poff"] + 1] = 1.0 / R
    params[els[kcap]["poff"]] = Cv
    freqs = np.logspace(1, 9, nf)
    t = api.Topology(base.num_nodes, els, p0)
    r = api.ac(t, params, freqs)
    x, st = host(r.x), host(r.status)
    assert (st == 0).all()
    out = base.node("out") - 1
    w = 2 * np.pi * freqs[:, None]
    H = 1.0 / (1.0 + 1j * w * R[None, :] * Cv[None, :])
    assert np.max(np.abs(x[:, out, :] - H)) <= 1e-12
    for b in range(0, B, 37):
        e, xo, _ = oracle.ac(base.num_nodes, els, np.ascontiguousarray(params[:, b]), freqs)
        assert e == 0 and np.array_equal(x[:, :, b], xo), b


# ---------------------------------------------------------------------------- backends
@pytest.mark.parametrize("name", ["bridge", "diode_fwd", "mos1_op", "mos1_cmos_vtc", "bjt_ce_op", "mos2_subvt_dc"])
def test_op_goldens_forced_sparse(name):
    """Options.Backend = SparseLU on circuits Auto would give the dense LU:
    the GP path against the oracle's sparse restatement, same status and
    iterations, solutions within 1e-9 (bit for bit where linear)."""
    c = gr.load_deck(name).circuit
    els, p = c.flatten()
    o = api.Options()
    o.backend = 2
    t = api.Topology(c.num_nodes, els, p)
    r = api.op(t, np.repeat(p[:, None], 8, axis=1), opts=o)
    x, st, it = host(r.x), host(r.status), host(r.iters)
    e0, x0, info = oracle.op(c.num_nodes, els, p, backend=2)
    assert (st == e0).all() and (it == info["iters"]).all()
    assert relerr(x, np.repeat(x0[:, None], 8, axis=1)).max() <= 1e-9
    if all(e.kind not in (6, 7, 8) for e in c.elements):
        assert (x == x0[:, None]).all()


def test_rectifier_batch_forced_dense_matches_oracle():
    """Options.Backend = DenseLU on the 35-unknown rectifier (Auto would pick
    the sparse LU): lu_dense (solver/dense.go:80-178) in the persistent
    kernel against the oracle's Dense, same status and step counts,
    solutions within 1e-9."""
    b = W.rectifier_rc_ladders(64)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    o = api.Options()
    o.backend = 1
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.tran(t, b.params, spec, opts=o, lane_spec=b.lane_spec)
    ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    xo, so, ao, ro, lo = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params, b.n, ospec, threads=8,
                                           lane_spec=b.lane_spec, opts=oracle.options(backend=1))
    st, acc, rej = host(r.status), host(r.accepted), host(r.rejected)
    assert (st == so).all() and (st == 0).all()
    same = (acc == ao) & (rej == ro)
    assert same.mean() >= 0.99, same.mean()
    assert relerr(host(r.x), xo)[same].max() <= 1e-9
    assert r.stats["patterns"] == 0  # the dense LU needs no ordering


def test_ac_forced_sparse_matches_oracle():
    """AC with Options.Backend = SparseLU on the BJT amplifier (Auto: CDense):
    the complex GP path (CSparse) against the oracle's, within 1e-9."""
    from kronspice_amd import netlist
    deck = gr.load_deck("bjt_ce_ac")
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "ac"][0].args
    freqs = np.array(netlist.ac_frequencies(a["variation"], a["points"], a["fstart"], a["fstop"]))
    els, p = c.flatten()
    e0, xo, mro = oracle.ac(c.num_nodes, els, p, freqs, opts=oracle.options(backend=2))
    assert e0 == 0
    o = api.Options()
    o.backend = 2
    t = api.Topology(c.num_nodes, els, p)
    r = api.ac(t, np.repeat(p[:, None], 4, axis=1), freqs, opts=o)
    x, st, pts = host(r.x), host(r.status), host(r.points)
    assert (st == 0).all() and (pts == len(freqs)).all()
    for k in range(4):
        assert np.max(np.abs(x[:, :, k] - xo)) <= 1e-9 * np.max(np.abs(xo)), k
