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
