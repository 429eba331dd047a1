"""Runs the reference's golden netlists (tests/golden/reference) through a
backend and compares with the golden CSVs using the reference's tolerances
(internal/golden/golden.go:23-27, :186, :292-295).

A backend is an object with
  op(circuit) -> (status, x, currents)                       (analysis.OP)
  tran(circuit, spec) -> (status, times, xs, currents, maxrel)   (analysis.Tran)
  ac(circuit, freqs) -> (status, X [nf, n])                    (analysis.AC)
"""
from __future__ import annotations

import math
import os
from typing import Callable, Dict, List

import numpy as np

from kronspice_amd import netlist
from kronspice_amd.circuit import Circuit

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "reference")

ABS_V, ABS_I, REL = 1e-6, 1e-12, 1e-3      # golden.go:23-27
TRAN_FULL_SCALE = 2e-3                      # golden.go:186
AC_REL, AC_PHASE_DEG = 1e-6, 1e-4           # golden.go:292-295


def names() -> List[str]:
    return sorted(f[:-4] for f in os.listdir(os.path.join(REF, "netlists")) if f.endswith(".cir"))


def load_deck(name: str) -> netlist.Deck:
    with open(os.path.join(REF, "netlists", name + ".cir")) as f:
        return netlist.parse(f.read())


def load_csv(path: str):
    rows, header = [], None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = [p.strip() for p in line.split(",")]
            if header is None:
                header = [p.lower() for p in parts]
                continue
            rows.append(parts)
    return header, rows


def values_op(c: Circuit, x: np.ndarray, cur: np.ndarray) -> Dict[str, float]:
    """golden.Values: v(node) for each node, i(name) for each element."""
    out = {}
    for i, nm in enumerate(c.node_names):
        if i == 0:
            continue
        out[f"v({nm})"] = float(x[i - 1])
    for k, e in enumerate(c.elements):
        out[f"i({e.name.lower()})"] = float(cur[k])
    return out


def compare_op(got: Dict[str, float], want: Dict[str, float]) -> List[str]:
    msgs = []
    for k in sorted(want):
        w = want[k]
        if k not in got:
            msgs.append(f"{k}: missing")
            continue
        ab = ABS_I if k.startswith("i(") else ABS_V
        tol = max(ab, REL * abs(w))
        if abs(got[k] - w) > tol:
            msgs.append(f"{k}: got {got[k]:.9g}, want {w:.9g} (tol {tol:.3g})")
    return msgs


def interp(xs, ys, t):
    return float(np.interp(t, xs, ys))


def series(c: Circuit, name: str, xs: np.ndarray, cur: np.ndarray):
    if name.startswith("v("):
        nd = c.node(name[2:-1])
        return xs[:, nd - 1] if nd > 0 else np.zeros(len(xs))
    k = c.element_index(name[2:-1])
    return cur[:, k]


def check(name: str, backend, ngspice: bool = False) -> List[str]:
    """Run one golden netlist through the backend; return mismatch messages."""
    deck = load_deck(name)
    c = deck.circuit
    path = os.path.join(REF, "ngspice" if ngspice else "golden", name + ".csv")
    header, rows = load_csv(path)
    kinds = [a.kind for a in deck.analyses]
    msgs: List[str] = []
    if header[0] == "name":
        st, x, cur = backend.op(c)
        if st != 0:
            return [f"status {st}"]
        want = {r[0].lower(): float(r[1]) for r in rows}
        return compare_op(values_op(c, x, cur), want)
    if header[0] == "sweep":
        a = [a for a in deck.analyses if a.kind == "dc"][0]
        src = a.args["source"]
        k = c.element_index(src)
        vals = np.array([[float(v) for v in r] for r in rows])
        cols = header[1:]
        for row in vals:
            c.elements[k].value = row[0]
            st, x, cur = backend.op(c)
            if st != 0:
                return [f"sweep {row[0]}: status {st}"]
            got = values_op(c, x, cur)
            for j, cn in enumerate(cols):
                w = row[j + 1]
                ab = ABS_I if cn.startswith("i(") else ABS_V
                tol = max(ab, REL * abs(w))
                if abs(got[cn] - w) > tol:
                    msgs.append(f"sweep {row[0]} {cn}: got {got[cn]:.9g} want {w:.9g}")
        return msgs
    if header[0] == "freq":
        a = [a for a in deck.analyses if a.kind == "ac"][0]
        freqs = netlist.ac_frequencies(a.args["variation"], a.args["points"], a.args["fstart"], a.args["fstop"])
        st, X = backend.ac(c, freqs)
        if st != 0:
            return [f"status {st}"]
        vals = np.array([[float(v) for v in r] for r in rows])
        for j, cn in enumerate(header[1:]):
            node = cn[3:-1]
            nd = c.node(node)
            for r in vals:
                fi = int(np.argmin(np.abs(np.asarray(freqs) - r[0])))
                if abs(freqs[fi] - r[0]) > 1e-9 * r[0]:
                    msgs.append(f"freq {r[0]} not in sweep")
                    continue
                ph = X[fi, nd - 1]
                if cn.startswith("vm("):
                    got, w = abs(ph), r[j + 1]
                    if abs(got - w) > AC_REL * abs(w) + 1e-15:
                        msgs.append(f"{cn} f={r[0]}: got {got:.12g} want {w:.12g}")
                elif cn.startswith("vp("):
                    got, w = math.degrees(math.atan2(ph.imag, ph.real)), r[j + 1]
                    d = abs(got - w)
                    d = min(d, abs(d - 360))
                    if d > AC_PHASE_DEG:
                        msgs.append(f"{cn} f={r[0]}: got {got:.9g} want {w:.9g}")
        return msgs
    # transient table
    a = [a for a in deck.analyses if a.kind == "tran"][0]
    st, times, xs, cur, maxrel = backend.tran(c, a.args)
    if st != 0:
        return [f"status {st}"]
    vals = np.array([[float(v) for v in r] for r in rows])
    for j, cn in enumerate(header[1:]):
        y = series(c, cn, xs, cur)
        ref = vals[:, j + 1]
        full = np.max(np.abs(ref))
        tol = max(TRAN_FULL_SCALE * full, ABS_V)
        if ngspice:
            t0, t1 = vals[0, 0], vals[-1, 0]
            m = (times >= t0) & (times <= t1)
            refi = np.interp(times[m], vals[:, 0], ref)
            d = np.abs(y[m] - refi)
            # bracketDistance relaxation at discontinuities (golden.go:240-262)
            idx = np.searchsorted(vals[:, 0], times[m])
            for q in np.nonzero(d > tol)[0]:
                i = idx[q]
                if 0 < i < len(ref):
                    lo, hi = min(ref[i - 1], ref[i]), max(ref[i - 1], ref[i])
                    v = y[m][q]
                    d[q] = min(d[q], lo - v if v < lo else (v - hi if v > hi else 0.0))
            worst = float(d.max()) if len(d) else 0.0
        else:
            worst = float(np.max(np.abs(np.interp(vals[:, 0], times, y) - ref)))
        if worst > tol:
            msgs.append(f"{cn}: max error {worst:.3g} (tol {tol:.3g})")
    if maxrel > 1e-9:
        msgs.append(f"ledger residual {maxrel:.3g}")
    return msgs
