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
