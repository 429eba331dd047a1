"""Launch-configuration autotuner for the batched analyses.

The persistent kernel can run each circuit in place in device memory or with
its working set staged in LDS, with one or several circuits per wavefront,
and compiled for 1, 2 or 4 resident wavefronts per SIMD (register budget).
Which is fastest depends on the circuit size (LDS footprint), the number of
circuits and the divergence between them, so the choice is
measured: each candidate runs a sample of the batch, and the fastest
candidate (circuits per second on the sample) is the configuration for the
whole batch. Results do not depend on the choice — every configuration runs
the same device code on the same values — only time does.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import api

# (staging, circuits per wavefront, resident wavefronts per SIMD) candidates.
CANDIDATES: Tuple[Tuple[int, int, int], ...] = (
    (0, 1, 1), (1, 1, 1), (1, 4, 1), (0, 4, 1),
    (0, 1, 2), (1, 1, 2), (1, 4, 2), (0, 4, 2),
    (0, 1, 4), (1, 1, 4), (0, 4, 4))


@dataclass
class Choice:
    staging: int
    circuits_per_wave: int
    waves_per_simd: int
    seconds: float
    timings: List[Tuple[int, int, int, float]]

    def options(self, base: Optional[api.Options] = None) -> api.Options:
        o = api.Options() if base is None else base
        o.staging = self.staging
        o.circuits_per_wave = self.circuits_per_wave
        o.waves_per_simd = self.waves_per_simd
        return o


def tune(run: Callable[[api.Options], object], candidates: Sequence[Tuple[int, int, int]] = CANDIDATES,
         base: Optional[api.Options] = None) -> Choice:
    """run(options) performs the analysis on the sample; the fastest candidate
    wins. Candidates the library refuses (KhError) are skipped."""
    import torch
    timings = []
    for staging, cpw, waves in candidates:
        o = api.Options() if base is None else api.Options.from_buffer_copy(base)
        o.staging, o.circuits_per_wave, o.waves_per_simd = staging, cpw, waves
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(o)
            torch.cuda.synchronize()
            timings.append((staging, cpw, waves, time.perf_counter() - t0))
        except api.KhError:
            continue
    if not timings:
        raise RuntimeError("no launch configuration ran")
    s, c, w, t = min(timings, key=lambda r: r[3])
    return Choice(s, c, w, t, timings)


def sample_columns(B: int, k: int) -> np.ndarray:
    """k circuits spread over the batch (the sample the candidates run)."""
    return np.linspace(0, B - 1, min(k, B)).astype(int)
