"""The wavefront-per-circuit kernel (kh_options.circuit_wave = 1, k_wave).

Bars: the same as the lane-per-circuit kernel's — the stamped values are
summed per slot in the reference's Add order and the LU follows
solver/sparse.go operation for operation, so every result is bit-identical to
the lane-per-circuit kernel's (which the other GPU tests compare with the
oracle), and linear circuits are bit-identical to the oracle.
"""
import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle
from tests import golden_runner as gr

pytestmark = [pytest.mark.gpu, pytest.mark.skip(
    reason="k_wave is switched off on the GPU: its first device run (the 1,024-ladder OP) did not finish "
           "within 180 s and the cause is not known; tests/test_wave_host.py checks its device code on the CPU")]


def host(t):
    return t.cpu().numpy()


def wave_opts(**kw):
    return api.Options(circuit_wave=1, **kw)


def test_wave_ladders_op_bitexact():
    b = W.resistor_ladders(1024)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.op(t, b.params, opts=wave_opts())
    xo, so, io = oracle.batch_op(b.circuit.num_nodes, b.elements, b.params, b.n, threads=8)
    assert (host(r.status) == 0).all() and (so == 0).all()
    assert np.array_equal(host(r.x), xo)


@pytest.mark.parametrize("name", [n for n in gr.names() if "tran" not in n and "_ac" not in n])
@pytest.mark.parametrize("backend", [0, 2])
def test_wave_op_goldens_equal_lane_kernel(name, backend):
    deck = gr.load_deck(name)
    c = deck.circuit
    els, p = c.flatten()
    t = api.Topology(c.num_nodes, els, p)
    P = np.repeat(p[:, None], 8, axis=1)
    rl = api.op(t, P, opts=api.Options(backend=backend), polish=True)
    rw = api.op(t, P, opts=wave_opts(backend=backend), polish=True)
    assert np.array_equal(host(rw.status), host(rl.status))
    assert np.array_equal(host(rw.iters), host(rl.iters))
    assert np.array_equal(host(rw.x), host(rl.x))


@pytest.mark.parametrize("name", [n for n in gr.names() if "tran" in n or n in ("rc_step", "rl_step", "rlc_step",
                                                                                 "rectifier_rc")])
def test_wave_tran_goldens_equal_lane_kernel(name):
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "tran"][0].args
    els, p = c.flatten(a["step"], a["stop"])
    spec = api.TranSpec(step=a["step"], stop=a["stop"], start=a["start"], max_step=a["max_step"],
                        uic=1 if a["uic"] else 0)
    t = api.Topology(c.num_nodes, els, p)
    rl = api.tran(t, p[:, None], spec, opts=api.Options(backend=2))
    rw = api.tran(t, p[:, None], spec, opts=wave_opts(backend=2))
    for f in ("status", "accepted", "rejected", "x", "ledger"):
        assert np.array_equal(host(getattr(rw, f)), host(getattr(rl, f))), f


def test_wave_rectifier_batch_equals_lane_kernel_and_oracle():
    b = W.rectifier_rc_ladders(256)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    rl = api.tran(t, b.params, spec, lane_spec=b.lane_spec)
    rw = api.tran(t, b.params, spec, lane_spec=b.lane_spec, opts=wave_opts())
    for f in ("status", "accepted", "rejected", "x", "ledger"):
        assert np.array_equal(host(getattr(rw, f)), host(getattr(rl, f))), f
    ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    xo, so, ao, ro, lo = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params, b.n, ospec, threads=8,
                                           lane_spec=b.lane_spec)
    assert (host(rw.status) == so).all()
    same = (host(rw.accepted) == ao) & (host(rw.rejected) == ro)
    assert same.mean() >= 0.99
    x = host(rw.x)
    rel = np.max(np.abs(x - xo), axis=0) / np.max(np.abs(xo), axis=0)
    assert rel[same].max() <= 1e-9


def test_wave_uneven_mosfet_chains_equal_lane_kernel():
    """An uneven batch (MOSFET chains of 8 patterns, 20-2,000 nodes) in one
    run: operating points and short transients bit-identical to the
    lane-per-circuit kernel's, forced sparse."""
    bs = [W.mosfet_chain_batch(p, 6, 3, steps=40) for p in (0, 17, 33, 58, 91, 120, 151, 199)]
    topos = [api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0)) for b in bs]
    spec = api.TranSpec(step=bs[0].spec["step"], stop=bs[0].spec["stop"], max_step=bs[0].spec["max_step"])
    lanes = [b.lane_spec for b in bs]
    outl = api.run_groups(topos, [b.params for b in bs], spec=spec, opts=api.Options(backend=2), lane_specs=lanes)
    outw = api.run_groups(topos, [b.params for b in bs], spec=spec, opts=wave_opts(backend=2), lane_specs=lanes)
    for gl, gw in zip(outl[0], outw[0]):
        for f in ("status", "accepted", "rejected", "x"):
            assert np.array_equal(host(getattr(gw, f)), host(getattr(gl, f))), f


@pytest.mark.parametrize("method,itl1", [("gmin", 4), ("source", 4), ("ptran", 2)])
def test_wave_fallback_chain_equals_lane_kernel(method, itl1):
    from tests.test_dc_fallbacks import METHOD_FLAGS, hard_batch
    c, els, P = hard_batch(64)
    t = api.Topology(c.num_nodes, els, P[:, 0])
    kw = dict(itl1=itl1, backend=2, **METHOD_FLAGS[method])
    rl = api.op(t, P, opts=api.Options(**kw), polish=True)
    rw = api.op(t, P, opts=wave_opts(**kw), polish=True)
    for f in ("status", "iters", "x"):
        assert np.array_equal(host(getattr(rw, f)), host(getattr(rl, f))), f
