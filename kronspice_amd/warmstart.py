"""The learned Newton warm start for one topology.

A small network (kh_mlp_*: genes -> 256 -> 256 -> node voltages) is trained on
operating points the solver itself produces for random gene draws, then its
predictions are handed to the batched OP as Options.NodeSet guesses
(kh_group.d_nodeset, analysis/newton.go:221-243): Newton starts at the guess
with every device's limiting seeded from it, and the usual fallbacks (gmin
and source stepping, pseudo-transient) remain for circuits the guess does
not bring home. The guess changes where Newton starts, not what it accepts:
both runs are polished to round-off, so a warm-started circuit's answer is
the cold one's to within the polish tolerance.

The workload is the two-stage BJT amplifier (workloads.bjt_amplifiers: 200
nodes, 60 genes).
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import api
from . import workloads as W


@dataclass
class Data:
    genes: np.ndarray      # [B, 60] network inputs in [-1, 1]
    volts: np.ndarray      # [B, nodes] DC node voltages (nodes 1..numNodes-1)
    status: np.ndarray     # [B] OP status
    iters: np.ndarray      # [B] Newton iterations of the cold start
    x: np.ndarray          # [n, B] polished solutions


def solve(batch: W.Batch, nodeset=None) -> tuple:
    """Polished OP of every circuit (cold, or from the given node voltages [nodes, B])."""
    t = api.Topology(batch.circuit.num_nodes, batch.elements, batch.circuit_params(0))
    r = api.op(t, batch.params, polish=True, nodeset=nodeset)
    return (r.x.cpu().numpy(), r.status.cpu().numpy(), r.iters.cpu().numpy())


def make_data(B: int, seed: int = 4, first: int = 0) -> Data:
    batch, genes = W.bjt_amplifiers(B, seed=seed, first_circuit=first)
    x, st, it = solve(batch)
    nodes = batch.circuit.num_nodes - 1
    return Data(genes.T.astype(np.float32).copy(), x[:nodes].T.copy(), st, it, x)


class WarmStart:
    """Genes -> node-voltage network with the per-node normalisation of its targets."""

    def __init__(self, n_genes: int, nodes: int, hidden: int = 256, seed: int = 1):
        self.net = api.MLP(n_genes, nodes, hidden=hidden, seed=seed)
        self.mean = np.zeros(nodes)
        self.std = np.ones(nodes)

    def fit(self, data: Data, steps: int = 3000, batch: int = 1024, lr: float = 2e-3, seed: int = 0,
            log_every: int = 0) -> list:
        """Adam on the mean squared error of normalised voltages (circuits whose
        operating point failed are left out). Returns the loss history."""
        torch = api._torch()
        ok = data.status == 0
        v = data.volts[ok]
        self.mean = v.mean(axis=0)
        self.std = np.maximum(v.std(axis=0), 1e-6)
        x = torch.from_numpy(data.genes[ok]).cuda()
        t = torch.from_numpy(((v - self.mean) / self.std).astype(np.float32)).cuda()
        n = x.shape[0]
        rng = np.random.default_rng(seed)
        hist = []
        for s in range(steps):
            idx = torch.from_numpy(rng.integers(0, n, size=min(batch, n))).cuda()
            loss = self.net.gradients(x[idx], t[idx])
            self.net.adam_step(lr * (0.1 ** (s / steps)))  # exponential decay to lr/10
            hist.append(loss)
            if log_every and s % log_every == 0:
                print(f"step {s}: loss {loss:.4g}", flush=True)
        return hist

    def predict(self, genes: np.ndarray) -> np.ndarray:
        """genes [B, 60] -> node voltages [nodes, B] (the NodeSet layout)."""
        torch = api._torch()
        y = self.net.forward(torch.from_numpy(np.ascontiguousarray(genes, np.float32)).cuda()).cpu().numpy()
        return (y.astype(np.float64) * self.std + self.mean).T.copy()


def evaluate(ws: WarmStart, B: int, seed: int = 4, first: int = 1 << 20) -> dict:
    """Cold against warm-started OP on B fresh circuits (not in the training set):
    Newton iterations, statuses, and the largest relative difference between the
    two polished solutions."""
    batch, genes = W.bjt_amplifiers(B, seed=seed, first_circuit=first)
    t0 = time.perf_counter()
    xc, sc, ic = solve(batch)
    t1 = time.perf_counter()
    guess = ws.predict(genes.T)
    t2 = time.perf_counter()
    xw, sw, iw = solve(batch, nodeset=guess)
    t3 = time.perf_counter()
    both = (sc == 0) & (sw == 0)
    den = np.maximum(np.max(np.abs(xc), axis=0), 1e-300)
    rel = np.max(np.abs(xw - xc), axis=0) / den
    nodes = batch.circuit.num_nodes - 1
    gerr = np.max(np.abs(guess - xc[:nodes]), axis=0)
    return {
        "circuits": B,
        "cold_status_ok": float(np.mean(sc == 0)),
        "warm_status_ok": float(np.mean(sw == 0)),
        "cold_iters_mean": float(np.mean(ic[both])),
        "warm_iters_mean": float(np.mean(iw[both])),
        "iters_saved_fraction": float(1 - np.sum(iw[both]) / max(np.sum(ic[both]), 1)),
        "max_rel_diff_polished": float(np.max(rel[both])) if both.any() else None,
        "guess_err_volts_median": float(np.median(gerr[both])) if both.any() else None,
        "cold_seconds": t1 - t0,
        "predict_seconds": t2 - t1,
        "warm_seconds": t3 - t2,
    }


def main(argv: Optional[list] = None) -> int:
    ap = argparse.ArgumentParser(description="train and evaluate the learned warm start")
    ap.add_argument("--train", type=int, default=65536)
    ap.add_argument("--eval", type=int, default=65536)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--batch", type=int, default=1024, help="training mini-batch")
    ap.add_argument("--hidden", type=int, default=256, help="hidden width (a multiple of 64)")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    data = make_data(a.train)
    t1 = time.perf_counter()
    ws = WarmStart(data.genes.shape[1], data.volts.shape[1], hidden=a.hidden)
    hist = ws.fit(data, steps=a.steps, batch=a.batch, lr=a.lr)
    t2 = time.perf_counter()
    rep = evaluate(ws, a.eval)
    rep.update(train_circuits=a.train, data_seconds=t1 - t0, train_seconds=t2 - t1, train_steps=a.steps, batch=a.batch, hidden=a.hidden,
               loss_first=hist[0], loss_last=float(np.mean(hist[-50:])))
    print(json.dumps(rep))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
