"""The batched caller of a whole generation (kh_generation_run): candidates of
several topologies in shuffled order, grouped by the library and run as one
uneven batch, with each candidate's results returned at its own index.

Bars: the same results, bit for bit, as running each topology's candidates
as their own batch (the grouping changes where a circuit runs, not what it
computes), for the DC operating point and the transient, and the oracle's
answers on a sample within 1e-9 (absolute 1e-15 where an operating point is
round-off noise).
"""
import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu


def mixed_generation(seed=0):
    ladders = W.resistor_ladders(24)
    rect = W.rectifier_rc_ladders(24, steps=200)
    mos = W.mosfet_chain_batch(7, 12, steps=100)
    cands = []
    for b in (ladders, rect, mos):
        for c in range(b.B):
            lane = None if b.lane_spec is None else np.ascontiguousarray(b.lane_spec[:, c])
            cands.append((b, c, (b.circuit.num_nodes, b.elements, b.circuit_params(c), lane)))
    order = np.random.default_rng(seed).permutation(len(cands))
    return [cands[i] for i in order]


def test_generation_op_equals_per_topology_batches():
    gen = mixed_generation()
    xs, st, it, _, _, stats = api.run_generation([c for _, _, c in gen])
    assert stats["patterns"] >= 3
    for b in {id(b): b for b, _, _ in gen}.values():
        t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
        r = api.op(t, b.params)
        x0, s0, i0 = r.x.cpu().numpy(), r.status.cpu().numpy(), r.iters.cpu().numpy()
        for k, (bb, c, _) in enumerate(gen):
            if bb is b:
                assert st[k] == s0[c] and it[k] == i0[c]
                assert np.array_equal(xs[k], x0[:, c])
    for k, (b, c, _) in enumerate(gen[:12]):
        e, xo, info = oracle.op(b.circuit.num_nodes, b.elements, b.circuit_params(c))
        assert e == st[k]
        # a rectifier's operating point (its sine source at 0 V) is round-off
        # noise of order 1e-18 V, which the last ulp of exp moves: there the
        # bar is absolute, 1e-15 V
        assert np.max(np.abs(xs[k] - xo)) <= 1e-9 * max(np.max(np.abs(xo)), 1e-6)


def test_generation_tran_equals_per_topology_batches():
    gen = [g for g in mixed_generation(1) if g[2][3] is not None]  # the transient workloads carry time scales
    b0 = gen[0][0]
    spec = api.TranSpec(step=b0.spec["step"], stop=b0.spec["stop"], max_step=b0.spec["max_step"])
    xs, st, it, acc, rej, stats = api.run_generation([c for _, _, c in gen], spec=spec)
    for b in {id(b): b for b, _, _ in gen}.values():
        t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
        r = api.tran(t, b.params, spec, lane_spec=b.lane_spec)
        x0, s0 = r.x.cpu().numpy(), r.status.cpu().numpy()
        a0, r0 = r.accepted.cpu().numpy(), r.rejected.cpu().numpy()
        for k, (bb, c, _) in enumerate(gen):
            if bb is b:
                assert st[k] == s0[c] and acc[k] == a0[c] and rej[k] == r0[c]
                assert np.array_equal(xs[k], x0[:, c])


def diode_clamp(rs, v, seed):
    """One element table whose diode has series resistance rs: RS > 0 gives
    the diode an internal node (device/diode.go), RS = 0 does not."""
    from kronspice_amd import circuit as kc
    c = (kc.Builder("clamp").vsource("V1", "a", "0", v).resistor("R1", "a", "b", 1e3 + 37 * seed)
         .diode("D1", "b", "0", "dm").resistor("R2", "b", "0", 5e3).model("dm", "d", {"is": 1e-14, "rs": rs})
         .build())
    els, p = c.flatten()
    return c.num_nodes, els, p, None


def test_generation_separates_internal_node_layouts():
    """Candidates with one element table but RS = 0 and RS > 0 diodes have
    different unknown layouts (3 and 4 unknowns): the generation groups them
    apart, each candidate's x has its own length, and each agrees with its
    own batch and with the oracle."""
    cands = [diode_clamp(0.0 if k % 3 else 2.5, 2.0 + 0.25 * k, k) for k in range(24)]
    sizes = [api.lib().kh_candidate_size(api.C.byref(_cand(c))) for c in cands]
    assert set(sizes) == {3, 4}
    xs, st, it, _, _, stats = api.run_generation(cands)
    assert (st == 0).all()
    for k, (nn, els, p, _) in enumerate(cands):
        assert len(xs[k]) == sizes[k]
        e, xo, info = oracle.op(nn, els, p)
        assert e == 0 and info["iters"] == it[k]
        assert np.max(np.abs(xs[k] - xo)) <= 1e-9 * np.max(np.abs(xo))


def test_batch_rejects_mixed_internal_node_layouts():
    nn, els, p0, _ = diode_clamp(2.5, 3.0, 0)
    _, _, p1, _ = diode_clamp(0.0, 3.0, 1)
    t = api.Topology(nn, els, p0)
    with pytest.raises(api.KhError) as e:
        api.op(t, np.stack([p0, p1], axis=1))
    assert e.value.status == api.BAD_ARGUMENT


def _cand(c):
    nn, els, prm, _ = c
    k = api.Candidate()
    k.num_nodes, k.n_elems, k.nparams = nn, len(els), len(prm)
    k.elems = np.ascontiguousarray(els).ctypes.data
    k.params = np.ascontiguousarray(prm, np.float64).ctypes.data
    k.lane_spec = None
    _cand.keep = (els, prm)
    return k
