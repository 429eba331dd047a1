"""The autotuner's choice among probe timings (kronspice_amd/autotune.py)."""
from kronspice_amd import autotune


def test_fastest_probe_wins_outside_the_tie_band():
    t = [(1, 1, 4, 0, 0.150), (1, 16, 1, 0, 0.202), (0, 16, 1, 0, 0.285)]
    assert autotune.pick(t) == t[0]


def test_ties_go_to_more_circuits_per_wavefront_then_fewer_waves():
    t = [(1, 1, 4, 0, 0.20168), (1, 16, 1, 0, 0.20204), (1, 16, 2, 0, 0.20190), (1, 8, 1, 0, 0.20603)]
    assert autotune.pick(t) == (1, 16, 1, 0, 0.20204)


def test_ties_keep_the_fp64_factorization():
    t = [(1, 16, 1, 1, 0.2010), (1, 16, 1, 0, 0.2015)]
    assert autotune.pick(t) == (1, 16, 1, 0, 0.2015)
