"""Compiled schedules on the GPU (kronspice_amd/csrc/kh_sched.h, k_ac_sched):
AC systems of one pattern factored in lockstep along the structure of a probe
system's factorization, each lane checking the reference's pivot rule
(solver/sparse.go:254-277) on its own values and handing a departure to the
per-lane LU.

Bars: a sweep with schedules is bit-identical to the same sweep with every
system factored alone (Options.no_schedule = 1), which the other AC tests
compare with the oracle — solutions, statuses, points, ledgers, element
currents; on the RLC mesh (whose inductor branch columns pivot off the
diagonal at low frequencies and on it at high ones) most systems run in
lockstep along a few schedules; on circuits whose pivot rows differ between
value draws, the departures are refused and redone, and the results still
agree bit for bit.
"""
import numpy as np
import pytest

from kronspice_amd import api, circuit as kc, netlist, workloads as W
from oracle import oracle
from tests import golden_runner as gr

pytestmark = pytest.mark.gpu


def host(t):
    return t.cpu().numpy()


def both(t, params, freqs, **kw):
    r1 = api.ac(t, params, freqs, opts=api.Options(**kw), currents=True)
    r0 = api.ac(t, params, freqs, opts=api.Options(no_schedule=1, **kw), currents=True)
    return r1, r0


def assert_same(r1, r0):
    assert np.array_equal(host(r1.status), host(r0.status))
    assert np.array_equal(host(r1.points), host(r0.points))
    assert np.array_equal(host(r1.x), host(r0.x))
    assert np.array_equal(host(r1.ledger), host(r0.ledger))
    assert np.array_equal(host(r1.currents), host(r0.currents))
    assert np.array_equal(host(r1.ledgers), host(r0.ledgers))


def test_sched_rlc_mesh_all_lockstep_and_identical():
    b = W.rlc_grid(16, 16, draws=4)
    f = W.rlc_grid_frequencies(64)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r1, r0 = both(t, b.params, f)
    assert (host(r1.status) == 0).all()
    assert_same(r1, r0)
    s1, s0 = r1.stats, r0.stats
    print(f"mesh 16x16: lockstep {s1['sched_solves']}, redone {s1['sched_redos']}")
    assert s1["sched_solves"] + s1["sched_redos"] == b.B * len(f)
    assert s1["sched_solves"] >= b.B * len(f) // 2
    assert s0["sched_solves"] == 0 and s0["sched_redos"] == 0


def test_sched_rlc_mesh_matches_oracle_bit_for_bit():
    b = W.rlc_grid(12, 12, draws=3)
    f = W.rlc_grid_frequencies(48)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r = api.ac(t, b.params, f)
    assert r.stats["sched_solves"] > 0
    x = host(r.x)
    for d in range(b.B):
        e, xo, _ = oracle.ac(b.circuit.num_nodes, b.elements, b.circuit_params(d), f)
        assert e == 0
        assert np.array_equal(x[:, :, d], xo), d


def perturbed(c, els, p, B, seed):
    """B copies of the deck's parameters with every R, C and L value scaled by
    log-U(1/3, 3) per copy (copy 0 unchanged)."""
    rng = np.random.default_rng(seed)
    P = np.repeat(p[:, None], B, axis=1).copy()
    for k, e in enumerate(c.elements):
        po = els[k]["poff"]
        s = np.exp(rng.uniform(np.log(1 / 3), np.log(3), B))
        s[0] = 1.0
        if e.kind == kc.R:
            P[po] = p[po] * s
            P[po + 1] = 1.0 / P[po]
        elif e.kind in (kc.C, kc.L):
            P[po] = p[po] * s
    return P


@pytest.mark.parametrize("name", ["rc_lowpass_ac", "rlc_bandpass_ac", "mos1_cs_ac", "bjt_ce_ac", "mos3_cs_ac",
                                  "param_rc_ac"])
def test_sched_forced_sparse_decks_identical(name):
    """The AC goldens under forced sparse (their size is below the dense
    threshold), 64 value draws each: whatever the draws' pivots do, the
    schedule's results equal the per-lane ones bit for bit."""
    deck = gr.load_deck(name)
    c = deck.circuit
    a = [a for a in deck.analyses if a.kind == "ac"][0].args
    freqs = np.array(netlist.ac_frequencies(a["variation"], a["points"], a["fstart"], a["fstop"]))
    els, p = c.flatten()
    P = perturbed(c, els, p, 64, 11)
    t = api.Topology(c.num_nodes, els, p)
    r1, r0 = both(t, P, freqs, backend=2)
    assert_same(r1, r0)
    s = r1.stats
    print(f"{name}: lockstep {s['sched_solves']}, redone {s['sched_redos']}")
    assert s["sched_solves"] + s["sched_redos"] > 0


def test_sched_pivot_departures_are_redone():
    """Draws whose pivots differ: in every odd draw the capacitors of every
    fifth node are scaled down by 1e9, which moves those nodes' diagonal
    magnitudes against their neighbours' and so where the diagonal preference
    holds; such systems depart from the schedules of the even draws' pivot
    sequences (and get their own, or are redone alone); all results equal the
    per-lane sweep's."""
    b = W.rlc_grid(6, 6, draws=8)
    f = W.rlc_grid_frequencies(32)
    P = b.params.copy()
    caps = [k for k, e in enumerate(b.circuit.elements) if e.kind == kc.C]
    for d in range(1, b.B, 2):
        for k in caps[::5]:
            P[b.elements[k]["poff"], d] *= 1e-9
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    r1, r0 = both(t, P, f)
    assert_same(r1, r0)
    s = r1.stats
    print(f"departures: lockstep {s['sched_solves']}, redone {s['sched_redos']}")
    assert s["sched_solves"] > 0
