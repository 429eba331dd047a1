"""Benchmark workloads at their full shape on the GPU, against the oracle.

* The two-stage BJT amplifier (200 nodes, 60 value genes over ±1 decade):
  4,096 circuits' polished DC operating points (analysis/op.go:243-271) from
  the batched OP, every status checked, 256 of them against the oracle at
  1e-9 relative with the same Newton iteration count.
* The 64×64 RLC grid mesh (n = 20,224 complex unknowns): 2 value draws at 4
  of the 1,024 log-spaced frequencies of the sweep (1 kHz, the two inner
  thirds and 10 GHz), every unknown bit-identical to the oracle's CSparse
  sweep (solver/sparse.go:495-531: the circuit is linear and the complex
  arithmetic follows Go's operation order).
"""
import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_bjt_amplifier_polished_op_matches_oracle():
    b, _ = W.bjt_amplifiers(4096)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.op(t, b.params, polish=True)
    x, st, it = r.x.cpu().numpy(), r.status.cpu().numpy(), r.iters.cpu().numpy()
    assert (st == 0).all(), np.unique(st, return_counts=True)
    for c in range(0, b.B, 16):
        e0, x0, it0 = oracle.op_polished(b.circuit.num_nodes, b.elements, b.circuit_params(c))
        assert e0 == 0, c
        assert it[c] == it0, (c, it[c], it0)
        rel = np.max(np.abs(x[:, c] - x0)) / np.max(np.abs(x0))
        assert rel <= 1e-9, (c, rel)


def test_rlc_grid_64x64_matches_oracle_bit_for_bit():
    b = W.rlc_grid(64, 64, draws=2)
    assert b.n == 20224
    f = W.rlc_grid_frequencies(1024)[[0, 341, 682, 1023]]
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.ac(t, b.params, f)
    x, st, pts = r.x.cpu().numpy(), r.status.cpu().numpy(), r.points.cpu().numpy()
    assert (st == 0).all() and (pts == len(f)).all()
    for d in range(b.B):
        e, xo, _ = oracle.ac(b.circuit.num_nodes, b.elements, b.circuit_params(d), f)
        assert e == 0
        assert np.array_equal(x[:, :, d], xo), d
