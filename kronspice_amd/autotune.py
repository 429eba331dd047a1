"""Launch-configuration autotuner for the batched analyses.

The persistent kernel can run each circuit in place in device memory or with
its working set staged in LDS, with one or several circuits per wavefront,
and compiled for 1, 2 or 4 resident wavefronts per SIMD (register budget);
its sparse LU runs in FP64 or as FP32 factors with FP64 refinement.
Which is fastest depends on the circuit size (LDS footprint), the number of
circuits and the divergence between them, so the choice is
measured: each candidate runs a probe — the whole batch (so the number of
circuits in flight, which decides how well latencies overlap, is the real
one) over a short stretch of the analysis — and the fastest candidate is the
configuration for the full run. The launch configuration is tuned in FP64,
then each other LU precision is probed at the chosen configuration. Results
do not depend on the launch configuration — every configuration runs the
same device code on the same values — only time does; the mixed precision
LU agrees with the FP64 path to round-off (its pivot rows are certified
equal to FP64's) rather than bit for bit.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import api

# (staging, circuits per wavefront, resident wavefronts per SIMD) candidates.
CANDIDATES: Tuple[Tuple[int, int, int], ...] = tuple(
    (s, c, w) for w in (1, 2, 4) for c in (1, 2, 4, 8, 16, 32) for s in (0, 1) if not (s == 1 and c > 16))


@dataclass
class Choice:
    staging: int
    circuits_per_wave: int
    waves_per_simd: int
    precision: int
    seconds: float
    timings: List[Tuple[int, int, int, int, float]]  # (staging, cpw, waves, precision, seconds)

    def options(self, base: Optional[api.Options] = None) -> api.Options:
        o = api.Options() if base is None else base
        o.staging = self.staging
        o.circuits_per_wave = self.circuits_per_wave
        o.waves_per_simd = self.waves_per_simd
        o.precision = self.precision
        return o


def tune(run: Callable[[api.Options], object], candidates: Sequence[Tuple[int, int, int]] = CANDIDATES,
         base: Optional[api.Options] = None, precisions: Sequence[int] = (0, 1)) -> Choice:
    """run(options) performs the analysis on the sample; the fastest candidate
    wins. Candidates the library refuses (KhError) are skipped."""
    import torch

    def probe(staging, cpw, waves, prec):
        o = api.Options() if base is None else api.Options.from_buffer_copy(base)
        o.staging, o.circuits_per_wave, o.waves_per_simd, o.precision = staging, cpw, waves, prec
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(o)
            torch.cuda.synchronize()
            timings.append((staging, cpw, waves, prec, time.perf_counter() - t0))
        except api.KhError:
            pass

    timings: List[Tuple[int, int, int, int, float]] = []
    for staging, cpw, waves in candidates:
        probe(staging, cpw, waves, precisions[0])
    if not timings:
        raise RuntimeError("no launch configuration ran")
    s, c, w, _, _ = pick(timings)
    for prec in precisions[1:]:
        probe(s, c, w, prec)
    s, c, w, pr, t = pick(timings)
    return Choice(s, c, w, pr, t, timings)


TIE = 0.01  # probe times within 1 % of the fastest are a tie (run-to-run spread of a probe)


def pick(timings):
    """The fastest probe; among probes within TIE of it, the one with the most
    circuits per wavefront, then the fewest wavefronts per SIMD, then the
    lowest precision flag: the same work then touches fewer cache lines per
    solve (the staged rectifier batch measures 4.0 KB of HBM traffic per
    solve at 16 circuits per wavefront against 53.7 KB at one), so ties do
    not flip between runs onto a configuration that moves more bytes."""
    t_best = min(r[4] for r in timings)
    close = [r for r in timings if r[4] <= t_best * (1 + TIE)]
    return min(close, key=lambda r: (-r[1], r[2], r[3], r[4]))


def short_lane_spec(lane_spec: np.ndarray, fraction: float) -> np.ndarray:
    """Per-circuit transient specs [4, B] (step, stop, max step, min step)
    cut to the first `fraction` of each circuit's interval: the probe."""
    s = np.array(lane_spec, dtype=np.float64, copy=True)
    s[1] *= fraction
    s[2] = np.minimum(s[2], s[1] / 50)
    return s
