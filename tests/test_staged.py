"""LDS-staged persistent kernel (kh_options.staging = 1) and several circuits
per wavefront: the same parity bars as test_gpu_parity.py."""
import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle
from tests import golden_runner as gr
from tests.test_gpu_parity import host, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cpw", [1, 4])
@pytest.mark.parametrize("name", ["bridge", "diode_fwd", "mos1_op", "zener_dc", "mos1_cmos_vtc", "bjt_ce_op"])
def test_op_goldens_staged(name, cpw):
    """Staged runs give the in-place runs' results bit for bit (the same
    device code on copies of the same arrays), hence the oracle's."""
    c = gr.load_deck(name).circuit
    els, p = c.flatten()
    t = api.Topology(c.num_nodes, els, p)
    B = 64
    params = np.repeat(p[:, None], B, axis=1)
    r = api.op(t, params, opts=api.Options(staging=1, circuits_per_wave=cpw))
    r0 = api.op(t, params)
    e0, x0, info = oracle.op(c.num_nodes, els, p)
    assert (host(r.status) == e0).all() and (host(r.iters) == info["iters"]).all()
    assert np.array_equal(host(r.x), host(r0.x)), np.max(np.abs(host(r.x) - host(r0.x)))


@pytest.mark.parametrize("cpw", [1, 4])
@pytest.mark.parametrize("name", ["rc_step", "diode_charge_tran", "mos1_inv_tran", "rectifier_rc", "bjt_switch_tran"])
def test_tran_goldens_staged(name, cpw):
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "tran"][0].args
    els, p = c.flatten(a["step"], a["stop"])
    spec = api.TranSpec(step=a["step"], stop=a["stop"], start=a["start"], max_step=a["max_step"],
                        uic=1 if a["uic"] else 0)
    t = api.Topology(c.num_nodes, els, p)
    B = 8
    r = api.tran(t, np.repeat(p[:, None], B, axis=1), spec, opts=api.Options(staging=1, circuits_per_wave=cpw))
    ospec = oracle.KhTranSpec(step=a["step"], stop=a["stop"], start=a["start"], max_step=a["max_step"],
                              uic=1 if a["uic"] else 0)
    e0, x0, info = oracle.tran(c.num_nodes, els, p, ospec)
    assert (host(r.status) == e0).all()
    assert (host(r.accepted) == info["accepted"]).all() and (host(r.rejected) == info["rejected"]).all()
    assert relerr(host(r.x), np.repeat(x0[:, None], B, axis=1)).max() <= 1e-9


@pytest.mark.parametrize("cpw", [1, 2, 4])
def test_rectifier_batch_staged(cpw):
    b = W.rectifier_rc_ladders(256)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.tran(t, b.params, spec, opts=api.Options(staging=1, circuits_per_wave=cpw), lane_spec=b.lane_spec)
    ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    xo, so, ao, ro, lo = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params, b.n, ospec, threads=8,
                                           lane_spec=b.lane_spec)
    st, acc, rej = host(r.status), host(r.accepted), host(r.rejected)
    assert (st == so).all()
    same = (acc == ao) & (rej == ro)
    assert same.mean() >= 0.99, same.mean()
    assert relerr(host(r.x), xo)[same].max() <= 1e-9


@pytest.mark.parametrize("waves,staging,cpw", [(2, 0, 1), (4, 0, 1), (2, 1, 4), (4, 1, 1)])
def test_register_budgets_bit_identical(waves, staging, cpw):
    """The persistent kernel compiled for 2 or 4 resident wavefronts per SIMD
    (fewer registers, spills) computes exactly what the unconstrained build
    computes."""
    b = W.rectifier_rc_ladders(256)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r0 = api.tran(t, b.params, spec, lane_spec=b.lane_spec)
    r = api.tran(t, b.params, spec, opts=api.Options(staging=staging, circuits_per_wave=cpw, waves_per_simd=waves),
                 lane_spec=b.lane_spec)
    assert np.array_equal(host(r.status), host(r0.status))
    assert np.array_equal(host(r.accepted), host(r0.accepted)) and np.array_equal(host(r.rejected), host(r0.rejected))
    assert np.array_equal(host(r.x), host(r0.x))


def tran_fields_equal(r, r0):
    for f in ("status", "accepted", "rejected", "x", "ledger"):
        assert np.array_equal(host(getattr(r, f)), host(getattr(r0, f))), f


def test_rectifier_4096_partially_staged_bit_identical():
    """4,096 circuits at 16 per wavefront make 256 workgroups, one per compute
    unit, so a circuit's LDS share is 160 KiB / 16 = 10 KiB, below its whole
    working set: the most used arrays are staged and the rest stays in device
    memory (staged_partial). Every result equals the in-place run's bit for
    bit."""
    b = W.rectifier_rc_ladders(4096)
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.tran(t, b.params, spec, opts=api.Options(staging=1, circuits_per_wave=16), lane_spec=b.lane_spec)
    assert r.stats["staged_partial"] == 1 and 0 < r.stats["staged_bytes"] <= 10240, r.stats
    r0 = api.tran(t, b.params, spec, opts=api.Options(circuits_per_wave=16), lane_spec=b.lane_spec)
    assert r0.stats["staged_bytes"] == 0
    tran_fields_equal(r, r0)


@pytest.mark.parametrize("cpw,waves", [(4, 1), (1, 4)])
def test_uneven_mosfet_chains_partially_staged_bit_identical(cpw, waves):
    """MOSFET chains of 8 patterns (20 to 2,000 nodes) in one launch: the large
    circuits stage only part of their arrays. Operating points and short
    transients equal the in-place run's bit for bit."""
    bs = [W.mosfet_chain_batch(p, 64, 3, steps=40) for p in (0, 17, 33, 58, 91, 120, 151, 199)]
    topos = [api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0)) for b in bs]
    spec = api.TranSpec(step=bs[0].spec["step"], stop=bs[0].spec["stop"], max_step=bs[0].spec["max_step"])
    lanes = [b.lane_spec for b in bs]
    kw = dict(backend=2, circuits_per_wave=cpw, waves_per_simd=waves)
    out0 = api.run_groups(topos, [b.params for b in bs], spec=spec, opts=api.Options(**kw), lane_specs=lanes)
    out1 = api.run_groups(topos, [b.params for b in bs], spec=spec, opts=api.Options(staging=1, **kw),
                          lane_specs=lanes)
    assert out1[1]["staged_partial"] >= 1, out1[1]
    for g0, g1 in zip(out0[0], out1[0]):
        for f in ("status", "accepted", "rejected", "x"):
            assert np.array_equal(host(getattr(g1, f)), host(getattr(g0, f))), f


def test_bjt_amplifier_op_4096_partially_staged_bit_identical():
    """The 200-node two-stage BJT amplifier's polished operating point over
    4,096 circuits, staged at 16 per wavefront (partially: its working set
    exceeds the 10 KiB share), equals the in-place run bit for bit."""
    b, _ = W.bjt_amplifiers(4096)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.op(t, b.params, opts=api.Options(staging=1, circuits_per_wave=16), polish=True)
    assert r.stats["staged_partial"] == 1, r.stats
    r0 = api.op(t, b.params, opts=api.Options(circuits_per_wave=16), polish=True)
    for f in ("status", "iters", "x"):
        assert np.array_equal(host(getattr(r, f)), host(getattr(r0, f))), f
