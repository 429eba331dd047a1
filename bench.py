"""Benchmark of the MI355X solve path: circuits solved per second.

Workload (one "step" = one batched analysis over the whole batch, inputs
already resident in HBM):
  rectifier_rc_ladder_tran — 4096 half-wave rectifier + 32-section RC ladder
  circuits of one topology (sine amplitude U(1,20) V, f log-U(50 Hz, 1 MHz);
  diode IS log-U(1e-15,1e-12) A, N U(1,2); R log-U(10,10k) ohm, C log-U(1n,10u) F),
  each run for 4 periods of its own source with max step = stop/2000:
  trapezoidal integration with LTE step control, Newton to reltol 1e-3,
  abstol 1e-12, the reference's DC operating point first (analysis.Tran).

Forms: the plain FP64 GPU path (in place in HBM, one circuit per wavefront)
runs the batch once; the chosen form is the launch configuration the
autotuner (kronspice_amd/autotune.py) picks on a sample of the batch
(LDS staging and circuits per wavefront), and it is what `value` times.
Both compute the same results bit for bit.

Prints one JSON line. `roofline` is for the dominant kernel (k_persist, the
persistent factor+solve+stamp kernel; every other kernel of the run is a
layout change or a fingerprint at the run's edges): algorithmic bytes
8*(nnzA + 2n + P) per executed solve, summed over the run, divided by the
summed HIP-event duration of its launches (events recorded on the stream the
library launches on). `cpu_baseline` times the FP64
CPU restatement of the reference (oracle/, one worker thread per core, one
task per circuit) on a bounded sample of the same batch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "circuits solved/sec at ≤1e-9 rel err vs FP64 oracle: chosen form vs plain GPU"
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_persist_rectifier4096_hbm.md")


def measured_traffic_per_solve(staging: int, cpw: int, waves: int):
    """HBM bytes per executed solve of k_persist on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary for the same launch
    configuration (None when that configuration was not measured)."""
    key = f"traffic_bytes_per_solve[staging={staging},cpw={cpw},waves={waves}]:"
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--parity-sample", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpw", type=int, default=0, help="circuits per wavefront (0: autotuned)")
    ap.add_argument("--staging", type=int, default=-1, help="0 in place, 1 LDS staged (-1: autotuned)")
    ap.add_argument("--waves", type=int, default=0, help="resident wavefronts per SIMD (0: autotuned)")
    ap.add_argument("--tune-fraction", type=float, default=0.05,
                    help="fraction of each circuit's interval the autotuner's probe simulates")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain FP64 form")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the learned warm start's measurement (two-stage BJT amplifier DC OP)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); circuits shard across GPUs with no "
              "exchange, so N GPUs would each run this workload on their own share.", file=sys.stderr)
        sys.exit(2)

    import torch
    from kronspice_amd import api, autotune, workloads as W

    b = W.rectifier_rc_ladders(args.batch)
    topo = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    spec = api.TranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    d_params = api.to_device(b.params)
    d_lane = api.to_device(b.lane_spec)
    torch.cuda.synchronize()

    # Plain FP64 GPU form: in place, one circuit per wavefront.
    plain_s = None
    if not args.no_plain:
        t0 = time.perf_counter()
        api.tran(topo, d_params, spec, opts=api.Options(staging=0, circuits_per_wave=1), lane_spec=d_lane)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0

    # Chosen form: the autotuner's pick on a sample of the batch.
    tuned = None
    if args.staging < 0 or args.cpw <= 0 or args.waves <= 0:
        sl = api.to_device(autotune.short_lane_spec(b.lane_spec, args.tune_fraction))
        tuned = autotune.tune(lambda o: api.tran(topo, d_params, spec, opts=o, lane_spec=sl))
        args.staging, args.cpw, args.waves = tuned.staging, tuned.circuits_per_wave, tuned.waves_per_simd
    opts = api.Options(circuits_per_wave=args.cpw, staging=args.staging, waves_per_simd=args.waves)

    res = None
    for _ in range(args.warmup):
        res = api.tran(topo, d_params, spec, opts=opts, lane_spec=d_lane)
    torch.cuda.synchronize()
    lu_s = ctl_s = 0.0
    lu_n = 0
    bytes_alg = 0
    solves = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = api.tran(topo, d_params, spec, opts=opts, lane_spec=d_lane)
        s = res.stats
        lu_s += s["lu_seconds"]
        ctl_s += s["control_seconds"]
        lu_n += s["lu_launches"]
        bytes_alg += s["bytes_algorithmic"]
        solves += s["solves"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = res.status.cpu().numpy()
    acc = res.accepted.cpu().numpy()
    x_gpu = res.x.cpu().numpy()
    ms_per_step = dt / args.steps * 1e3

    # Parity on a sample: circuits whose status is OK and whose final state is
    # within 1e-9 of the oracle count as solved.
    from oracle import oracle
    ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
    k = min(args.parity_sample, b.B)
    idx = np.linspace(0, b.B - 1, k).astype(int)
    threads = min(16, os.cpu_count() or 1)  # the host cores one GPU's share of the machine has
    xo, so, ao, ro, lo = oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params[:, idx], b.n, ospec,
                                           threads=threads, lane_spec=b.lane_spec[:, idx])
    rel = np.max(np.abs(x_gpu[:, idx] - xo), axis=0) / np.maximum(np.max(np.abs(xo), axis=0), 1e-300)
    ok = (st[idx] == 0) & (so == 0) & (rel <= 1e-9)
    solved_frac = float(ok.mean())
    value = b.B * solved_frac / (ms_per_step / 1e3)

    cpu = None
    if not args.no_cpu:
        m = min(args.cpu_sample, b.B)
        sidx = np.arange(m)
        tc = time.perf_counter()
        oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params[:, sidx], b.n, ospec, threads=threads,
                          lane_spec=b.lane_spec[:, sidx])
        tcpu = time.perf_counter() - tc
        cpu = {"value": m / tcpu, "unit": "circuits/s", "cores": threads, "kind": "port",
               "sample": f"first {m} circuits of the batch, FP64 C++ restatement of the reference's Go path, "
                         f"{threads} worker threads, one task per circuit ({tcpu:.2f} s)"}

    warm = None
    if not args.no_warm_start:
        # The learned Newton warm start (the metric's learned form) on its own
        # workload: a 200-node two-stage BJT amplifier, DC operating points of
        # fresh gene draws, cold against network-seeded Newton.
        from kronspice_amd import warmstart as WS
        tw = time.perf_counter()
        wdata = WS.make_data(65536)
        ws = WS.WarmStart(wdata.genes.shape[1], wdata.volts.shape[1])
        hist = ws.fit(wdata, steps=8000, lr=3e-3)
        trained = time.perf_counter() - tw
        rep = WS.evaluate(ws, 65536)
        warm = {"workload": "two-stage BJT amplifier, 200 nodes, 60 genes, DC OP (polished)",
                "train_circuits": 65536, "train_seconds": trained, "eval_circuits": 65536,
                "cold_iters_mean": rep["cold_iters_mean"], "warm_iters_mean": rep["warm_iters_mean"],
                "iters_saved_fraction": rep["iters_saved_fraction"],
                "cold_circuits_per_s": 65536 / rep["cold_seconds"],
                "warm_circuits_per_s": 65536 / (rep["predict_seconds"] + rep["warm_seconds"]),
                "status_ok": rep["warm_status_ok"], "max_rel_diff_polished": rep["max_rel_diff_polished"],
                "loss_last": float(np.mean(hist[-50:]))}

    achieved = bytes_alg / lu_s / 1e9 if lu_s > 0 else 0.0
    per_solve = measured_traffic_per_solve(args.staging, max(args.cpw, 1), max(args.waves, 1))
    traffic = per_solve * solves / max(lu_n, 1) if per_solve is not None else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "circuits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": "rectifier_rc_ladder_tran: half-wave rectifier + 32-section RC ladder, "
                               "4 source periods, max step stop/2000, trapezoidal + LTE, FP64 GP LU",
                   "circuits": b.B, "unknowns": b.n,
                   "form": f"chosen: autotuned launch (staging={args.staging}, circuits per wavefront={args.cpw}, "
                           f"wavefronts per SIMD={args.waves}), FP64 GP LU",
                   "parallelism": f"persistent kernel, {args.cpw} circuit(s) per wavefront, 1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_persist", "avg_launch_us": lu_s / max(lu_n, 1) * 1e6,
                     "bytes_per_launch": bytes_alg / max(lu_n, 1),
                     "traffic_source": "rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000 per solve "
                                       "(copy-kernel calibrated), profiles/k_persist_rectifier4096_hbm.md"},
        "cpu_baseline": cpu,
        "parity": {"sample": int(k), "solved_fraction": solved_frac, "max_rel_err": float(rel.max()),
                   "status_ok": float((st == 0).mean()), "accepted_steps_mean": float(acc.mean()),
                   "same_step_count": float((acc[idx] == ao).mean())},
        "forms": {"plain_fp64_gpu": ({"circuits_per_s": b.B / plain_s, "seconds": plain_s} if plain_s else None),
                  "chosen": {"circuits_per_s": b.B / (ms_per_step / 1e3), "staging": args.staging,
                             "circuits_per_wave": args.cpw, "waves_per_simd": args.waves},
                  "chosen_vs_plain": (plain_s / (ms_per_step / 1e3)) if plain_s else None,
                  "autotune": ({"probe": f"whole batch, first {args.tune_fraction:g} of each interval",
                                "seconds": {f"staging={a},cpw={c},waves={w}": t for a, c, w, t in tuned.timings}}
                               if tuned else None),
                  "learned_warm_start": warm},
        "detail": {"status_histogram": {int(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))},
                   "launches": res.stats["ticks"], "circuits_per_wave": args.cpw, "staging": args.staging,
                   "solves_per_step": solves / args.steps,
                   "cycles_lu_fraction": res.stats["cycles_lu"] / max(res.stats["cycles_lu"] + res.stats["cycles_control"], 1.0),
                   "control_ms_per_step": ctl_s / args.steps * 1e3, "lu_ms_per_step": lu_s / args.steps * 1e3},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
