"""Batched complex AC sweep of an RLC grid mesh (series R-L per edge, C per
node to ground, a unit AC current into one corner, the transfer impedance
read at the opposite corner): many value draws times many frequencies, each
(draw, frequency) one complex system of one sparsity pattern.

Bars: every unknown at every frequency bit-identical to the FP64 oracle's
complex GP sweep (the circuit is linear and complex arithmetic follows Go's
operation order, so no transcendental function enters the solve); the
network is reciprocal, so the transfer impedance with source and probe
swapped agrees to 1e-10 (a check that rests on physics, not on the oracle);
splitting the frequencies into chunks changes nothing.
"""
import os

import numpy as np
import pytest

from kronspice_amd import api, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu


def sweep(b, freqs):
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.ac(t, b.params, freqs)
    return r.x.cpu().numpy(), r.status.cpu().numpy(), r.points.cpu().numpy()


def test_rlc_grid_matches_oracle_bit_for_bit():
    b = W.rlc_grid(8, 8, draws=4)
    f = W.rlc_grid_frequencies(32)
    x, st, pts = sweep(b, f)
    assert (st == 0).all() and (pts == len(f)).all()
    for d in range(b.B):
        e, xo, _ = oracle.ac(b.circuit.num_nodes, b.elements, b.circuit_params(d), f)
        assert e == 0
        assert np.array_equal(x[:, :, d], xo), d


def test_rlc_grid_reciprocity():
    w = h = 8
    a = W.rlc_grid(w, h, draws=4)
    b = W.rlc_grid(w, h, draws=4, source_at_far_corner=True)
    f = W.rlc_grid_frequencies(32)
    xa, sa, _ = sweep(a, f)
    xb, sb, _ = sweep(b, f)
    assert (sa == 0).all() and (sb == 0).all()
    za = xa[:, a.circuit.node(f"g{w - 1}_{h - 1}") - 1, :]
    zb = xb[:, b.circuit.node("g0_0") - 1, :]
    big = np.abs(za) > 1e-250  # the mesh is a steep low-pass: far-corner values underflow at 10 GHz
    assert big.mean() > 0.5
    assert np.max(np.abs(za - zb)[big] / np.abs(za)[big]) <= 1e-10


def test_rlc_grid_chunked_sweep_identical():
    b = W.rlc_grid(8, 8, draws=4)
    f = W.rlc_grid_frequencies(32)
    x0, _, _ = sweep(b, f)
    os.environ["KH_AC_CHUNK_LANES"] = "20"  # 5 frequencies x 4 draws per chunk, 7 chunks
    try:
        x1, st, pts = sweep(b, f)
    finally:
        del os.environ["KH_AC_CHUNK_LANES"]
    assert (st == 0).all() and (pts == len(f)).all()
    assert np.array_equal(x0, x1)


def test_rlc_grid_mixed_precision_matches_complex128():
    """precision 1 on the AC sweep: complex64 factors + complex128 refinement,
    complex128 redo of whatever the mixed path cannot certify. Every system
    agrees with the complex128 sweep to 1e-10 of its largest unknown (the
    complex128 solution's own round-off on the mesh's worse-conditioned
    frequencies reaches ~1e-12)."""
    b = W.rlc_grid(8, 8, draws=4)
    f = W.rlc_grid_frequencies(32)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r0 = api.ac(t, b.params, f)
    r1 = api.ac(t, b.params, f, opts=api.Options(precision=1))
    x0, x1 = r0.x.cpu().numpy(), r1.x.cpu().numpy()
    assert (r1.status.cpu().numpy() == 0).all()
    den = np.max(np.abs(x0), axis=1)  # [nf, B]
    rel = np.max(np.abs(x1 - x0), axis=1) / np.maximum(den, 1e-300)
    s = r1.stats
    print(f"AC mixed: max rel {rel.max():.2e}, refine passes {s['refine_passes']}, redone {s['fp64_redos']}")
    assert rel.max() <= 1e-10
    assert s["refine_passes"] > 0
