"""Benchmark of the MI355X solve path: circuits solved per second.

Default workload (one "step" = one batched analysis over the whole batch,
inputs already resident in HBM):
  rectifier_rc_ladder_tran — 4096 half-wave rectifier + 32-section RC ladder
  circuits of one topology (sine amplitude U(1,20) V, f log-U(50 Hz, 1 MHz);
  diode IS log-U(1e-15,1e-12) A, N U(1,2); R log-U(10,10k) ohm,
  C log-U(1n,10u) F), each run for 4 periods of its own source with max step
  stop/2000: the DC operating point, then trapezoidal integration with LTE
  step control, Newton to reltol 1e-3, abstol 1e-12 (analysis.Tran).
  --workload mos_chains: 16384 uneven MOSFET-chain netlists over 200 sparsity
  patterns (20-2000 nodes), DC OP then 2 periods at max step stop/500, in one
  uneven batch (minutes per batch on the current kernel: run it on its own).

Forms: the plain FP64 GPU path (in place in HBM, one circuit per wavefront,
FP64 LU) runs the batch once; the chosen form is what the autotuner
(kronspice_amd/autotune.py) picks on a short probe of the whole batch —
LDS staging, circuits per wavefront, register budget, then FP64 LU against
FP32 factors + FP64 refinement — and it is what `value` times. The learned
Newton warm start (a per-topology network) is measured on its own workload,
the 200-node two-stage BJT amplifier's DC operating points, in
`forms.learned_warm_start`.

`value` counts the circuits solved per second: the batch size times the
fraction of the parity sample (the whole batch for one topology, one circuit
of every pattern for the uneven workload) whose status is OK and whose final
state is within 1e-9 of the FP64 oracle's, over the chosen form's time per
batch. The same sample, run by the FP64 C++ restatement of the reference
(oracle/, one task per circuit on the host's cores), gives `cpu_baseline`.

`roofline` is for the dominant kernel, k_persist (the persistent
factor+solve+stamp kernel; the run's other kernels are layout changes at its
edges): algorithmic bytes per executed solve — 8*(nnzA + 2n + P) for an
FP64 solve; 4*(nnzA + 2n) + 8*P plus 8*(nnzA + 2n) per further refinement
pass for a mixed one — summed over the run, divided by the summed HIP-event
duration of its launches (events on the stream the library launches on).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "circuits solved/sec at ≤1e-9 rel err vs FP64 oracle: chosen form vs plain GPU"
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILES = {"rectifier": os.path.join(ROOT, "profiles", "k_persist_rectifier4096_hbm.md"),
                    "mos_chains": os.path.join(ROOT, "profiles", "k_persist_mos_chains16384_hbm.md")}
HOST_THREADS = 16  # the host cores one GPU's share of the machine has


def measured_traffic_per_solve(workload: str, staging: int, cpw: int, waves: int, precision: int):
    """HBM bytes per executed solve of k_persist on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary for the same launch
    configuration (None when that configuration was not measured)."""
    keys = [f"traffic_bytes_per_solve[staging={staging},cpw={cpw},waves={waves},precision={precision}]:"]
    if precision == 0:
        keys.append(f"traffic_bytes_per_solve[staging={staging},cpw={cpw},waves={waves}]:")
    try:
        with open(TRAFFIC_PROFILES[workload]) as f:
            for line in f:
                for key in keys:
                    if line.startswith(key):
                        return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_ac_traffic(grid: int):
    """HBM bytes per k_ac_sched launch on the G x G mesh sweep, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    key = f"traffic_bytes_per_launch[grid={grid}]:"
    try:
        with open(os.path.join(ROOT, "profiles", f"k_ac_sched_rlc_grid{grid}_hbm.md")) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


class Work:
    """One batched workload: groups of circuits sharing a topology, run in
    one uneven batch through kh_batch_run."""

    def __init__(self, name, batches, description):
        from kronspice_amd import api
        import torch
        self.name, self.batches, self.description = name, batches, description
        self.B = sum(b.B for b in batches)
        self.topos = [api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0)) for b in batches]
        self.d_params = [api.to_device(b.params) for b in batches]
        self.d_lanes = [api.to_device(b.lane_spec) for b in batches]
        s = batches[0].spec
        self.spec = api.TranSpec(step=s["step"], stop=s["stop"], max_step=s["max_step"])
        torch.cuda.synchronize()

    def run(self, opts, lanes=None):
        from kronspice_amd import api
        return api.run_groups(self.topos, self.d_params, spec=self.spec, opts=opts,
                              lane_specs=lanes if lanes is not None else self.d_lanes)

    def probe_lanes(self, fraction):
        from kronspice_amd import api, autotune
        return [api.to_device(autotune.short_lane_spec(b.lane_spec, fraction)) for b in self.batches]

    def sample(self, per_group):
        """(group, column indices) of the parity / CPU-baseline sample."""
        return [(g, np.linspace(0, b.B - 1, min(per_group, b.B)).astype(int)) for g, b in enumerate(self.batches)]

    def oracle(self, sample, threads):
        """The FP64 oracle on the sample, one task per circuit on `threads`
        host threads; returns per-group results and the wall time."""
        from oracle import oracle

        def one(job):
            g, c = job
            b = self.batches[g]
            ospec = oracle.KhTranSpec(step=b.spec["step"], stop=b.spec["stop"], max_step=b.spec["max_step"])
            return oracle.batch_tran(b.circuit.num_nodes, b.elements, b.params[:, [c]], b.n, ospec, threads=1,
                                     lane_spec=b.lane_spec[:, [c]])
        jobs = [(g, int(c)) for g, cols in sample for c in cols]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            outs = list(ex.map(one, jobs))
        return dict(zip(jobs, outs)), time.perf_counter() - t0


def make_work(name, batch):
    from kronspice_amd import workloads as W
    if name == "rectifier":
        b = W.rectifier_rc_ladders(batch or 4096)
        return Work("rectifier", [b], "rectifier_rc_ladder_tran: half-wave rectifier + 32-section RC ladder, "
                                      "4 source periods, max step stop/2000, trapezoidal + LTE")
    bs = W.mosfet_chains(batch or 16384, patterns=200)
    return Work("mos_chains", bs, "mos_chains_uneven: 200 MOSFET-chain patterns (20-2000 nodes, 1-50 level-1 "
                                  "common-source stages), DC OP then 2 periods at max step stop/500, "
                                  "trapezoidal + LTE, one uneven batch")


def bench_ac(args, log):
    """The batched complex AC sweep of an RLC grid mesh: `--grid` G x G nodes,
    16 value draws x 1024 log-spaced frequencies (1 kHz - 10 GHz), each
    (draw, frequency) one complex system of one pattern. Forms: complex128 GP
    LU with every system alone on its lane (plain); complex128 GP LU in
    lockstep along compiled schedules (kh_sched.h: systems that share a pivot
    sequence share one structure; bit-identical); complex64 factors with
    complex128 refinement. Prints the same JSON fields as the transient
    workloads, in complex systems per second; with the scheduled form chosen,
    `roofline` is that of k_ac_sched (algorithmic bytes of the systems it
    solved over its summed launch time) and `traffic` its HBM bytes per launch
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    import torch
    from kronspice_amd import api, workloads as W
    from oracle import oracle
    G = args.grid
    b = W.rlc_grid(G, G, draws=16)
    freqs = W.rlc_grid_frequencies(1024)
    t = api.Topology(b.circuit.num_nodes, b.elements, b.circuit_params(0))
    d_params = api.to_device(b.params)
    torch.cuda.synchronize()
    systems = b.B * len(freqs)

    def timed(prec, sched):
        opts = api.Options(precision=prec, no_schedule=0 if sched else 1)
        r = None
        for _ in range(args.warmup):
            r = api.ac(t, d_params, freqs, opts=opts)
        torch.cuda.synchronize()
        tot = {"lu_seconds": 0.0, "bytes_algorithmic": 0.0, "lu_launches": 0.0, "refine_passes": 0.0, "fp64_redos": 0.0,
               "sched_solves": 0.0, "sched_redos": 0.0, "sched_probes": 0.0, "sched_rounds": 0.0,
               "sched_host_seconds": 0.0, "sched_seconds": 0.0}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = api.ac(t, d_params, freqs, opts=opts)
            for k in tot:
                tot[k] += r.stats[k]
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / args.steps, tot

    # Plain form: complex128 LU, every system alone on its lane. The others:
    # complex128 LU in lockstep along a compiled schedule (kh_sched.h), and
    # complex64 factors with complex128 refinement. The fastest is chosen.
    forms = {"plain": timed(0, False), "sched": timed(0, True), "c64": timed(1, False)}
    name = min(forms, key=lambda k: forms[k][1])
    prec = 1 if name == "c64" else 0
    r, dt, tot = forms[name]
    lu_s, bytes_alg, launches = tot["lu_seconds"], tot["bytes_algorithmic"], tot["lu_launches"]
    log(f"rlc grid {G}x{G}: n={b.n}, {systems} systems: complex128 alone {forms['plain'][1]:.3f} s, "
        f"complex128 scheduled {forms['sched'][1]:.3f} s, complex64+refinement {forms['c64'][1]:.3f} s")
    x = r.x.cpu().numpy()
    st = r.status.cpu().numpy()
    # Parity and CPU baseline: the oracle's sweep of every draw at 16 of the frequencies.
    fi = np.linspace(0, len(freqs) - 1, 16).astype(int)
    threads = min(HOST_THREADS, os.cpu_count() or 1)

    def one(d):
        return oracle.ac(b.circuit.num_nodes, b.elements, b.circuit_params(d), freqs[fi])
    tc = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        ref = list(ex.map(one, range(b.B)))
    tcpu = time.perf_counter() - tc
    ok, exact, rels = [], [], []
    for d, (e, xo, _) in enumerate(ref):
        for k, f in enumerate(fi):
            g = x[f, :, d]
            den = max(np.max(np.abs(xo[k])), 1e-300)
            rel = float(np.max(np.abs(g - xo[k])) / den)
            rels.append(rel)
            exact.append(bool(np.array_equal(g, xo[k])))
            ok.append(bool(e == 0 and st[d] == 0 and rel <= 1e-9))
    solved = float(np.mean(ok))
    achieved = bytes_alg / lu_s / 1e9 if lu_s > 0 else 0.0
    kernel, launches_k, secs_k, bytes_k = "k_ac_solve", launches, lu_s, bytes_alg
    traffic = None
    if name == "sched" and tot["sched_seconds"] > 0:
        # k_ac_sched: the systems it solved, 16*(nnzA + 2n) bytes each
        per_sys = bytes_alg / (systems * args.steps)
        kernel, launches_k, secs_k = "k_ac_sched", tot["sched_rounds"], tot["sched_seconds"]
        bytes_k = tot["sched_solves"] * per_sys
        achieved = bytes_k / secs_k / 1e9
        traffic = measured_ac_traffic(G)
    out = {
        "metric": METRIC, "value": systems * solved / dt, "unit": "complex systems/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128" if prec == 0 else "c64 LU + c128 refinement",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": f"rlc_grid_ac: {G}x{G} RLC mesh (series R-L per edge, C per node), unit AC current "
                               f"at one corner, 16 draws x 1024 frequencies 1 kHz-10 GHz",
                   "unknowns": b.n, "systems": systems,
                   "form": "chosen: " + {"plain": "complex128 GP LU, each system alone on its lane",
                                         "sched": "complex128 GP LU in lockstep along a compiled schedule, "
                                                  "lane per system",
                                         "c64": "complex64 GP LU + complex128 refinement, lane per system"}[name],
                   "parallelism": "lane per (draw, frequency) system, 1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "avg_launch_us": secs_k / max(launches_k, 1) * 1e6,
                     "bytes_per_launch": bytes_k / max(launches_k, 1)},
        "cpu_baseline": {"value": len(ok) / tcpu, "unit": "complex systems/s", "cores": threads, "kind": "port",
                         "sample": f"{b.B} draws x 16 frequencies, FP64 C++ restatement (one task per draw, its "
                                   f"operating point included) ({tcpu:.2f} s)"},
        "parity": {"sample": len(ok), "solved_fraction": solved, "bit_identical_fraction": float(np.mean(exact)),
                   "max_rel_err": float(max(rels))},
        "forms": {"plain_c128_gpu": {"systems_per_s": systems / forms["plain"][1], "seconds": forms["plain"][1]},
                  "c128_scheduled": {"systems_per_s": systems / forms["sched"][1], "seconds": forms["sched"][1],
                                     "lockstep_systems": forms["sched"][2]["sched_solves"] / args.steps,
                                     "redone_alone": forms["sched"][2]["sched_redos"] / args.steps,
                                     "probes": forms["sched"][2]["sched_probes"] / args.steps,
                                     "rounds": forms["sched"][2]["sched_rounds"] / args.steps,
                                     "probe_host_seconds": forms["sched"][2]["sched_host_seconds"] / args.steps,
                                     "kernel_seconds": forms["sched"][2]["lu_seconds"] / args.steps},
                  "c64_refined": {"systems_per_s": systems / forms["c64"][1], "seconds": forms["c64"][1],
                                  "refine_passes": forms["c64"][2]["refine_passes"] / args.steps,
                                  "c128_redos": forms["c64"][2]["fp64_redos"] / args.steps},
                  "chosen": name, "chosen_vs_plain": forms["plain"][1] / dt},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["rectifier", "mos_chains", "rlc_grid"], default="rectifier")
    ap.add_argument("--grid", type=int, default=16, help="rlc_grid: nodes per side")
    ap.add_argument("--batch", type=int, default=0, help="circuits (0: the workload's size)")
    ap.add_argument("--sample-per-group", type=int, default=0,
                    help="parity / CPU-baseline sample per group (0: 1 per pattern, the whole batch for one topology)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpw", type=int, default=0, help="circuits per wavefront (0: autotuned)")
    ap.add_argument("--staging", type=int, default=-1, help="0 in place, 1 LDS staged (-1: autotuned)")
    ap.add_argument("--waves", type=int, default=0, help="resident wavefronts per SIMD (0: autotuned)")
    ap.add_argument("--precision", type=int, default=-1, help="0 FP64 LU, 1 FP32 + FP64 refinement (-1: autotuned)")
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
    from kronspice_amd import api, autotune

    def log(msg):
        print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

    if args.workload == "rlc_grid":
        bench_ac(args, log)
        return

    work = make_work(args.workload, args.batch)
    log(f"{work.name}: {work.B} circuits in {len(work.batches)} groups")

    # Plain FP64 GPU form: in place, one circuit per wavefront, FP64 LU; the
    # same untimed warmup runs and timed steps as the chosen form.
    plain_s = plain_stats = None
    if not args.no_plain:
        plain = api.Options(staging=0, circuits_per_wave=1, waves_per_simd=1, precision=0)
        for _ in range(args.warmup):
            work.run(plain)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, plain_stats = work.run(plain)
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t0) / args.steps
        log(f"plain FP64 form: {plain_s:.2f} s per batch ({args.warmup} warmup, {args.steps} timed)")

    # Chosen form: the autotuner's pick on a short probe of the whole batch.
    tuned = None
    if args.staging < 0 or args.cpw <= 0 or args.waves <= 0 or args.precision < 0:
        lanes = work.probe_lanes(args.tune_fraction)
        tuned = autotune.tune(lambda o: work.run(o, lanes))
        args.staging, args.cpw, args.waves, args.precision = (tuned.staging, tuned.circuits_per_wave,
                                                              tuned.waves_per_simd, tuned.precision)
        log(f"autotuned: staging={args.staging} cpw={args.cpw} waves={args.waves} precision={args.precision}")
    opts = api.Options(circuits_per_wave=args.cpw, staging=args.staging, waves_per_simd=args.waves,
                       precision=args.precision)

    res = stats = None
    for _ in range(args.warmup):
        res, stats = work.run(opts)
    torch.cuda.synchronize()
    tot = {k: 0.0 for k in ("lu_seconds", "control_seconds", "lu_launches", "bytes_algorithmic", "solves",
                            "refine_passes", "fp64_redos")}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, stats = work.run(opts)
        for k in tot:
            tot[k] += stats[k]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    log(f"chosen form: {ms_per_step / 1e3:.2f} s per batch")

    # Parity on the sample against the FP64 oracle (whose time on the host's
    # cores is the CPU baseline).
    per = args.sample_per_group or (1 if len(work.batches) > 1 else work.B)
    sample = work.sample(per)
    threads = min(HOST_THREADS, os.cpu_count() or 1)
    ref, tcpu = work.oracle(sample, threads)
    log(f"oracle on {len(ref)} circuits: {tcpu:.2f} s")
    ok, rels, same = [], [], []
    for g, cols in sample:
        r = res[g]
        x = r.x.cpu().numpy()
        st = r.status.cpu().numpy()
        acc = r.accepted.cpu().numpy()
        for c in cols:
            xo, so, ao, ro, lo = ref[(g, int(c))]
            xo = xo[:, 0]
            rel = float(np.max(np.abs(x[:, c] - xo)) / max(np.max(np.abs(xo)), 1e-300))
            rels.append(rel)
            same.append(bool(acc[c] == ao[0]))
            ok.append(bool(st[c] == 0 and so[0] == 0 and rel <= 1e-9))
    solved_frac = float(np.mean(ok))
    value = work.B * solved_frac / (ms_per_step / 1e3)
    cpu = None
    if not args.no_cpu:
        cpu = {"value": len(ref) / tcpu, "unit": "circuits/s", "cores": threads, "kind": "port",
               "sample": f"{len(ref)} circuits ({min(per, work.B)} per topology group, every group) of the same batch, FP64 C++ "
                         f"restatement of the reference's Go path, {threads} worker threads, one task per circuit "
                         f"({tcpu:.2f} s)"}

    warm = None
    if not args.no_warm_start:
        # The learned Newton warm start on its own workload: a 200-node
        # two-stage BJT amplifier, DC operating points of fresh gene draws,
        # cold against network-seeded Newton.
        from kronspice_amd import warmstart as WS
        tw = time.perf_counter()
        wdata = WS.make_data(200000)  # 200k solver-generated operating points
        ws = WS.WarmStart(wdata.genes.shape[1], wdata.volts.shape[1])
        hist = ws.fit(wdata, steps=8000, batch=4096, lr=3e-3)
        del wdata
        trained = time.perf_counter() - tw
        rep = WS.evaluate(ws, 65536)
        warm = {"workload": "two-stage BJT amplifier, 200 nodes, 60 genes, DC OP (polished)",
                "train_circuits": 200000, "train_seconds": trained, "eval_circuits": 65536,
                "guess_err_volts_median": rep["guess_err_volts_median"],
                "cold_iters_mean": rep["cold_iters_mean"], "warm_iters_mean": rep["warm_iters_mean"],
                "iters_saved_fraction": rep["iters_saved_fraction"],
                "cold_circuits_per_s": 65536 / rep["cold_seconds"],
                "warm_circuits_per_s": 65536 / (rep["predict_seconds"] + rep["warm_seconds"]),
                "status_ok": rep["warm_status_ok"], "max_rel_diff_polished": rep["max_rel_diff_polished"],
                "loss_last": float(np.mean(hist[-50:]))}
        log(f"warm start: {warm['cold_iters_mean']:.2f} -> {warm['warm_iters_mean']:.2f} iterations")

    lu_s, lu_n = tot["lu_seconds"], max(tot["lu_launches"], 1)
    achieved = tot["bytes_algorithmic"] / lu_s / 1e9 if lu_s > 0 else 0.0
    per_solve = measured_traffic_per_solve(work.name, args.staging, max(args.cpw, 1), max(args.waves, 1),
                                           args.precision)
    traffic = per_solve * tot["solves"] / lu_n if per_solve is not None else None
    sts = np.concatenate([r.status.cpu().numpy() for r in res])
    accs = np.concatenate([r.accepted.cpu().numpy() for r in res])
    chosen_s = ms_per_step / 1e3
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
        "dtype": "f64" if args.precision == 0 else "f32 LU + f64 refinement",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": work.description, "circuits": work.B, "groups": len(work.batches),
                   "unknowns_max": max(b.n for b in work.batches),
                   "form": f"chosen: autotuned (staging={args.staging}, circuits per wavefront={args.cpw}, "
                           f"wavefronts per SIMD={args.waves}, precision={args.precision})",
                   "parallelism": f"persistent kernel, {args.cpw} circuit(s) per wavefront, 1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_persist", "avg_launch_us": lu_s / lu_n * 1e6,
                     "bytes_per_launch": tot["bytes_algorithmic"] / lu_n,
                     "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE per solve (copy-kernel calibrated), "
                                        + os.path.relpath(TRAFFIC_PROFILES[work.name], ROOT))
                     if per_solve is not None else "not measured for this launch configuration"},
        "cpu_baseline": cpu,
        "parity": {"sample": len(ok), "solved_fraction": solved_frac, "max_rel_err": float(max(rels)),
                   "same_step_count": float(np.mean(same)), "status_ok": float((sts == 0).mean()),
                   "accepted_steps_mean": float(accs.mean())},
        "forms": {"plain_fp64_gpu": ({"circuits_per_s": work.B / plain_s, "seconds": plain_s,
                                      "solves": plain_stats["solves"]} if plain_s else None),
                  "chosen": {"circuits_per_s": work.B / chosen_s, "seconds": chosen_s, "staging": args.staging,
                             "circuits_per_wave": args.cpw, "waves_per_simd": args.waves,
                             "precision": args.precision, "solves": tot["solves"] / args.steps,
                             "refine_passes": tot["refine_passes"] / args.steps,
                             "fp64_redos": tot["fp64_redos"] / args.steps},
                  "chosen_vs_plain": (plain_s / chosen_s) if plain_s else None,
                  "autotune": ({"probe": f"whole batch, first {args.tune_fraction:g} of each interval",
                                "seconds": {f"staging={a},cpw={c},waves={w},precision={p}": t
                                            for a, c, w, p, t in tuned.timings}} if tuned else None),
                  "learned_warm_start": warm},
        "detail": {"status_histogram": {int(k): int(v) for k, v in zip(*np.unique(sts, return_counts=True))},
                   "launches": stats["ticks"], "patterns": stats["patterns"],
                   "solves_per_step": tot["solves"] / args.steps,
                   "cycles_lu_fraction": stats["cycles_lu"] / max(stats["cycles_lu"] + stats["cycles_control"], 1.0),
                   "control_ms_per_step": tot["control_seconds"] / args.steps * 1e3,
                   "lu_ms_per_step": lu_s / args.steps * 1e3},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
