"""Mixed precision form of the solve path (kh_options.precision = 1: FP32
factors, FP64 iterative refinement, FP64 redo of anything the mixed path
cannot certify) against the FP64 path and the FP64 oracle.

Bars: pivot rows of every certified mixed factorization equal the FP64
factorization's row for row (the FP64 path is itself bit-exact with the
oracle, test_gpu_parity.py); refined solutions within 1e-12 of the FP64
solutions; analyses run in mixed precision end within 1e-9 of the oracle
(polished operating points, final transient states with the same step
counts on at least 99 % of circuits) with the same statuses.
"""
import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle
from tests import golden_runner as gr

pytestmark = pytest.mark.gpu


def host(t):
    return t.cpu().numpy()


def relerr(a, b):
    den = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return np.max(np.abs(a - b), axis=0) / den


def grid_mna(w, h, nsrc, B, rng):
    """Conductance grid (values log-uniform 1e-4..1 S, gmin 1e-12 per node)
    with nsrc voltage sources tying nodes to ground; entries in first-Add
    order, values [nnz, B]."""
    nn = w * h
    n = nn + nsrc
    ent = {}
    order = []

    def add(i, j, v):
        if (i, j) not in ent:
            ent[(i, j)] = np.zeros(B)
            order.append((i, j))
        ent[(i, j)] += v

    for y in range(h):
        for x in range(w):
            a = y * w + x
            for b in ([a + 1] if x + 1 < w else []) + ([a + w] if y + 1 < h else []):
                g = np.exp(np.log(1e-4) + rng.random(B) * np.log(1e4))
                add(a, a, g); add(b, b, g); add(a, b, -g); add(b, a, -g)
            add(a, a, np.full(B, 1e-12))
    for k in range(nsrc):
        node = (k * 7919) % nn
        add(node, nn + k, np.ones(B)); add(nn + k, node, np.ones(B))
    ri = np.array([i for i, _ in order], np.int32)
    ci = np.array([j for _, j in order], np.int32)
    vals = np.stack([ent[e] for e in order])
    rhs = np.zeros((n, B))
    rhs[nn:] = 1.0 + rng.random((nsrc, B))
    return n, ri, ci, vals, rhs


def ladder_mna(nn, B, rng):
    ri, ci = [], []
    for i in range(nn):
        ri.append(i); ci.append(i)
        if i + 1 < nn:
            ri += [i, i + 1]; ci += [i + 1, i]
    ri += [0, nn]; ci += [nn, 0]
    g = np.exp(np.log(1e-5) + rng.random((2 * nn, B)) * np.log(1e3))
    vals = []
    for i, j in zip(ri, ci):
        if i == j:
            vals.append(g[i] + (g[nn + i - 1] if i > 0 else 0) + (g[nn + i] if i + 1 < nn else 0))
        elif i == nn or j == nn:
            vals.append(np.ones(B))
        else:
            vals.append(-g[nn + min(i, j)])
    rhs = np.zeros((nn + 1, B)); rhs[nn] = 10.0
    return nn + 1, np.array(ri, np.int32), np.array(ci, np.int32), np.stack(vals), rhs


def run_both(n, ri, ci, vals, rhs):
    import torch
    B = vals.shape[1]
    lu = api.LUBatch(n, ri, ci, B)
    dv, db = api.to_device(vals), api.to_device(rhs)
    out = {}
    for mode in ("fp64", "mixed"):
        x = torch.zeros((n, B), dtype=torch.float64, device="cuda")
        st = torch.zeros(B, dtype=torch.int32, device="cuda")
        prow = torch.zeros((n, B), dtype=torch.int32, device="cuda")
        passes = torch.zeros(B, dtype=torch.int32, device="cuda")
        if mode == "fp64":
            lu.solve(dv, db, x, st, prow)
        else:
            lu.solve_mixed(dv, db, x, st, prow, passes)
        torch.cuda.synchronize()
        out[mode] = (host(x), host(st), host(prow), host(passes))
    return out


@pytest.mark.parametrize("case", ["ladder65", "grid12x9", "grid40x40"])
def test_lu_batch_mixed_pivots_equal_fp64(case):
    rng = np.random.default_rng(5)
    if case == "ladder65":
        sysm = ladder_mna(64, 1024, rng)
    elif case == "grid12x9":
        sysm = grid_mna(12, 9, 3, 512, rng)
    else:
        sysm = grid_mna(40, 40, 4, 64, rng)
    o = run_both(*sysm)
    x64, s64, p64, _ = o["fp64"]
    xm, sm, pm, passes = o["mixed"]
    assert (s64 == 0).all() and (sm == 0).all()
    cert = passes > 0
    print(f"{case}: certified {cert.mean():.4f}, passes mean {passes[cert].mean():.2f} max {passes.max()}")
    assert cert.mean() >= 0.99, cert.mean()
    assert np.array_equal(pm[:, cert], p64[:, cert])
    assert np.array_equal(xm[:, ~cert], x64[:, ~cert])  # redone in FP64: the FP64 path's own bits
    # The FP64 solution carries its own round-off (condition number times
    # 2^-53): grids of conductances over four decades leave a few 1e-12.
    assert relerr(xm, x64).max() <= (1e-12 if case == "ladder65" else 1e-10)


def test_lu_batch_mixed_random_unsymmetric():
    """Random unsymmetric systems (pivots among comparable values): whatever
    is certified has FP64's pivot rows, the rest is redone in FP64."""
    rng = np.random.default_rng(9)
    n, B = 40, 512
    mask = rng.random((n, n)) < 0.1
    np.fill_diagonal(mask, True)
    ri, ci = np.nonzero(mask)
    vals = (rng.random((len(ri), B)) - 0.5) * np.exp(rng.random((len(ri), B)) * 6 - 3)
    rhs = rng.random((n, B)) - 0.5
    o = run_both(n, ri.astype(np.int32), ci.astype(np.int32), vals, rhs)
    x64, s64, p64, _ = o["fp64"]
    xm, sm, pm, passes = o["mixed"]
    assert np.array_equal(sm, s64)
    ok = s64 == 0
    cert = (passes > 0) & ok
    assert cert.any() and (~cert & ok).any()
    assert np.array_equal(pm[:, cert], p64[:, cert])
    assert np.array_equal(xm[:, ~cert & ok], x64[:, ~cert & ok])


NONLINEAR_OP = [n for n in gr.names() if "tran" not in n and "_ac" not in n]


@pytest.mark.parametrize("name", NONLINEAR_OP)
def test_mixed_op_goldens_match_oracle(name):
    deck = gr.load_deck(name)
    c = deck.circuit
    els, p = c.flatten()
    t = api.Topology(c.num_nodes, els, p)
    B = 16
    r = api.op(t, np.repeat(p[:, None], B, axis=1), opts=api.Options(precision=1, backend=2), polish=True)
    e0, x0, _ = oracle.op_polished(c.num_nodes, els, p, opts=oracle.options(backend=2))
    st, x = host(r.status), host(r.x)
    assert (st == e0).all(), (st[0], e0)
    if e0 == 0:
        # Some decks' operating point is round-off noise (every unknown below
        # 1e-20, e.g. a source at 0 V): there the bar is absolute, 1e-18.
        den = max(np.max(np.abs(x0)), 1e-9)
        assert np.max(np.abs(x - x0[:, None])) / den <= 1e-9


def test_mixed_rectifier_batch_tran_matches_oracle():
    b = W.rectifier_rc_ladders(256)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.tran(t, b.params, spec, opts=api.Options(precision=1), lane_spec=b.lane_spec)
    ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    xo, so, ao, ro, lo = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params, b.n, ospec, threads=8,
                                           lane_spec=b.lane_spec)
    st, acc, rej = host(r.status), host(r.accepted), host(r.rejected)
    assert (st == so).all()
    same = (acc == ao) & (rej == ro)
    assert same.mean() >= 0.99, same.mean()
    assert relerr(host(r.x), xo)[same].max() <= 1e-9
    s = r.stats
    print(f"solves {s['solves']} refine passes {s['refine_passes']} fp64 redos {s['fp64_redos']}")
    assert s["refine_passes"] > 0
    assert s["fp64_redos"] <= 0.01 * s["solves"]
