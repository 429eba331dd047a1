"""Seeded synthetic batches for the solve path's benchmark configurations.

Every value is drawn from a counter-based generator (SplitMix64 of
seed, circuit index and field index), so any circuit regenerates in
isolation.  Log-uniform draws are exp(ln a + u (ln b - ln a)).

A batch is one topology (element table + parameter layout, include/kronhip.h)
and a parameter matrix params[P, B] (SoA: parameter-major, circuit-minor).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import circuit as kc

MASK64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finalizer on uint64 arrays (Steele, Lea, Flood 2014)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(MASK64)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(MASK64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(MASK64)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, circuits: np.ndarray, field: int) -> np.ndarray:
    """U[0,1) for (seed, circuit, field), 53-bit resolution."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed & MASK64) * np.uint64(0x100000001B3)) ^ (circuits.astype(np.uint64) << np.uint64(20)) \
            ^ np.uint64(field & 0xFFFFF)
        r = splitmix64(splitmix64(key))
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def log_uniform(seed, circuits, field, a, b):
    u = uniform(seed, circuits, field)
    return np.exp(math.log(a) + u * (math.log(b) - math.log(a)))


def lin_uniform(seed, circuits, field, a, b):
    return a + uniform(seed, circuits, field) * (b - a)


@dataclass
class Batch:
    """B circuits of one topology."""
    name: str
    circuit: kc.Circuit          # circuit 0 (template; values of circuit 0)
    elements: np.ndarray         # ELEMENT_DTYPE
    params: np.ndarray           # [P, B] float64
    n: int                       # MNA unknowns
    spec: Optional[dict] = None  # transient spec (uniform part)
    lane_spec: Optional[np.ndarray] = None  # [4, B]: step, stop, max_step, min_step per circuit

    @property
    def B(self) -> int:
        return self.params.shape[1]

    def circuit_params(self, c: int) -> np.ndarray:
        return np.ascontiguousarray(self.params[:, c])


def _template(c: kc.Circuit, B: int, step=0.0, stop=0.0):
    els, p0 = c.flatten(step, stop)
    params = np.repeat(p0[:, None], B, axis=1)
    return els, params


def _set_resistors(els, params, idx: List[int], values: np.ndarray):
    """values [len(idx), B]: R for each listed element (R and g = 1/R)."""
    for row, k in enumerate(idx):
        po = els[k]["poff"]
        params[po] = values[row]
        params[po + 1] = 1.0 / values[row]


# ---------------------------------------------------------------------------- config 1
def resistor_ladders(B: int = 1024, nodes: int = 64, seed: int = 1) -> Batch:
    """Plain DC OP: B resistor ladders of `nodes` nodes; 10 V at node 1, a shunt
    resistor at every node and a series resistor between consecutive nodes,
    R log-uniform in [100 ohm, 100 kohm]. MNA size nodes + 1 (one branch)."""
    b = kc.Builder("ladder").vsource("V1", "n1", "0", 10.0)
    ridx = []
    for i in range(1, nodes + 1):
        b.resistor(f"Rp{i}", f"n{i}", "0", 1e3)
        ridx.append(len(b.elems) - 1)
        if i < nodes:
            b.resistor(f"Rs{i}", f"n{i}", f"n{i + 1}", 1e3)
            ridx.append(len(b.elems) - 1)
    c = b.build()
    els, params = _template(c, B)
    cs = np.arange(B)
    vals = np.stack([log_uniform(seed, cs, f, 100.0, 1e5) for f in range(len(ridx))])
    _set_resistors(els, params, ridx, vals)
    _refresh(c, els, params)
    return Batch("resistor_ladder_dc_op", c, els, params, c.layout().n)


def _refresh(c: kc.Circuit, els, params):
    """Copy circuit 0's drawn values back into the template circuit."""
    for k, e in enumerate(c.elements):
        if e.kind == kc.R:
            e.value = float(params[els[k]["poff"], 0])
        elif e.kind in (kc.C, kc.L):
            e.value = float(params[els[k]["poff"], 0])


# ---------------------------------------------------------------------------- config 2
def rectifier_rc_ladders(B: int = 4096, sections: int = 32, seed: int = 2, periods: float = 4.0,
                         steps: int = 2000) -> Batch:
    """Batched nonlinear transient: 1 V-source SIN (amplitude U(1,20) V, f
    log-U(50 Hz, 1 MHz)) -> diode (IS log-U(1e-15,1e-12) A, N U(1,2)) -> an
    RC ladder of `sections` series-R / shunt-C sections (R log-U(10,10k) ohm,
    C log-U(1 nF, 10 uF)). Each circuit runs `periods` periods of its own
    frequency with max step = stop / steps (trapezoidal, LTE control)."""
    b = kc.Builder("rectifier_rc_ladder")
    b.vsource("V1", "in", "0", 0.0, wave=kc.Wave(kc.WAVE_SIN, (0.0, 10.0, 1e3)))
    b.diode("D1", "in", "o0", "dr")
    b.model("dr", "d", {"is": 1e-14, "n": 1.0})
    ridx, cidx = [], []
    for k in range(1, sections + 1):
        b.resistor(f"R{k}", f"o{k - 1}", f"o{k}", 1e3)
        ridx.append(len(b.elems) - 1)
        b.capacitor(f"C{k}", f"o{k}", "0", 1e-6)
        cidx.append(len(b.elems) - 1)
    c = b.build()
    cs = np.arange(B)
    amp = lin_uniform(seed, cs, 0, 1.0, 20.0)
    freq = log_uniform(seed, cs, 1, 50.0, 1e6)
    IS = log_uniform(seed, cs, 2, 1e-15, 1e-12)
    N = lin_uniform(seed, cs, 3, 1.0, 2.0)
    stop = periods / freq
    step = stop / steps
    els, params = _template(c, B, float(step[0]), float(stop[0]))
    # source: params 3..9 resolved SIN args (vo, va, f, td, theta, phase)
    pv = els[0]["poff"]
    params[pv + 3] = 0.0
    params[pv + 4] = amp
    params[pv + 5] = freq
    # diode: IS, N and the derived vte, vcrit (device/diode.go:116-117)
    pd = els[1]["poff"]
    vte = N * kc.BOLTZMANN_K * kc.TNOM_K / kc.CHARGE_Q
    params[pd + 0] = IS
    params[pd + 1] = N
    params[pd + 11] = vte
    params[pd + 12] = vte * np.log(vte / (math.sqrt(2) * IS))
    _set_resistors(els, params, ridx, np.stack([log_uniform(seed, cs, 10 + k, 10.0, 1e4) for k in range(sections)]))
    for k, ci in enumerate(cidx):
        params[els[ci]["poff"]] = log_uniform(seed, cs, 100 + k, 1e-9, 1e-5)
    _refresh(c, els, params)
    lane = np.stack([step, stop, step, 1e-12 * stop])
    return Batch("rectifier_rc_ladder_tran", c, els, params, c.layout().n,
                 spec=dict(step=float(step[0]), stop=float(stop[0]), start=0.0, max_step=float(step[0])),
                 lane_spec=lane)


def mosfet_chain_pattern(p: int, seed: int = 3) -> Tuple[int, List[int]]:
    """Stage count and per-stage RC-ladder lengths of pattern p: target node
    count log-uniform in [20, 2000], 1-50 stages of >= 3 nodes each."""
    pa = np.array([p])
    nodes = int(round(float(log_uniform(seed, pa, 0, 20.0, 2000.0)[0])))
    smax = max(1, min(50, (nodes - 2) // 3))
    stages = 1 + int(uniform(seed, pa, 1)[0] * smax)
    stages = min(stages, smax)
    rest = max(0, nodes - 2 - 2 * stages)
    base, extra = divmod(rest, stages)
    perm = np.argsort(uniform(seed, np.arange(stages) + 1000 * p, 2))
    ladders = [base + (1 if k in set(perm[:extra].tolist()) else 0) for k in range(stages)]
    return stages, ladders


def mosfet_chain_batch(p: int, B: int, seed: int = 3, steps: int = 500, first_circuit: int = 0) -> Batch:
    """One sparsity pattern of the uneven MOSFET-chain workload: a chain of
    common-source level-1 NMOS stages (KP log-U(2e-5,2e-4), VTO U(0.4,1.0) V,
    W/L U(1,100), LAMBDA 0.02), each biased by a resistive divider from a 5 V
    supply, AC-coupled from the previous stage, with a drain load resistor and
    capacitor and an RC ladder (R log-U(100,10k) ohm, C log-U(1p,1n) F) after
    the drain. The input is a small sine (U(10,100) mV, f log-U(1k,100k) Hz);
    each circuit runs the DC operating point and 2 input periods with max step
    stop/steps."""
    stages, ladders = mosfet_chain_pattern(p, seed)
    b = kc.Builder(f"mos_chain_{p}")
    b.vsource("VDD", "vdd", "0", 5.0)
    b.vsource("VIN", "in", "0", 0.0, wave=kc.Wave(kc.WAVE_SIN, (0.0, 0.05, 1e4)))
    idx: Dict[str, List[int]] = {k: [] for k in ("cc", "r1", "r2", "rd", "m", "cd", "rl", "cl")}
    prev = "in"
    for k in range(stages):
        g, d = f"g{k}", f"d{k}"
        b.capacitor(f"CC{k}", prev, g, 1e-7); idx["cc"].append(len(b.elems) - 1)
        b.resistor(f"RA{k}", "vdd", g, 1e5); idx["r1"].append(len(b.elems) - 1)
        b.resistor(f"RB{k}", g, "0", 1e5); idx["r2"].append(len(b.elems) - 1)
        b.resistor(f"RD{k}", "vdd", d, 1e4); idx["rd"].append(len(b.elems) - 1)
        b.mosfet(f"M{k}", d, g, "0", "0", "nch", {"l": 2e-6, "w": 20e-6}); idx["m"].append(len(b.elems) - 1)
        b.capacitor(f"CD{k}", d, "0", 1e-11); idx["cd"].append(len(b.elems) - 1)
        node = d
        for j in range(ladders[k]):
            nxt = f"l{k}_{j}"
            b.resistor(f"RL{k}_{j}", node, nxt, 1e3); idx["rl"].append(len(b.elems) - 1)
            b.capacitor(f"CL{k}_{j}", nxt, "0", 1e-10); idx["cl"].append(len(b.elems) - 1)
            node = nxt
        prev = node
    b.model("nch", "nmos", {"level": 1, "kp": 1e-4, "vto": 0.7, "lambda": 0.02})
    c = b.build()
    cs = np.arange(first_circuit, first_circuit + B) + 100000 * p
    f = log_uniform(seed, cs, 1, 1e3, 1e5)
    stop = 2.0 / f
    step = stop / steps
    els, params = _template(c, B, float(step[0]), float(stop[0]))
    pv = els[1]["poff"]
    params[pv + 4] = lin_uniform(seed, cs, 0, 0.01, 0.1)
    params[pv + 5] = f
    fld = 10
    vdd = 5.0
    for k in range(stages):
        kp = log_uniform(seed, cs, fld, 2e-5, 2e-4)
        vto = lin_uniform(seed, cs, fld + 1, 0.4, 1.0)
        wl = lin_uniform(seed, cs, fld + 2, 1.0, 100.0)
        beta = kp * wl
        pm = els[idx["m"][k]]["poff"]
        params[pm + 1] = beta
        params[pm + 2] = vto
        params[pm + 3] = vto
        vg = vto + lin_uniform(seed, cs, fld + 3, 0.1, 0.8)
        rtot = log_uniform(seed, cs, fld + 4, 2e4, 2e6)
        r2 = rtot * vg / vdd
        _set_resistors(els, params, [idx["r1"][k], idx["r2"][k]], np.stack([rtot - r2, r2]))
        idd = 0.5 * beta * (vg - vto) ** 2
        vd = lin_uniform(seed, cs, fld + 5, 1.5, 3.5)
        rd = np.clip((vdd - vd) / idd, 100.0, 1e6)
        _set_resistors(els, params, [idx["rd"][k]], rd[None, :])
        params[els[idx["cc"][k]]["poff"]] = log_uniform(seed, cs, fld + 6, 1e-8, 1e-6)
        params[els[idx["cd"][k]]["poff"]] = log_uniform(seed, cs, fld + 7, 1e-12, 1e-10)
        fld += 8
    for j, e in enumerate(idx["rl"]):
        _set_resistors(els, params, [e], log_uniform(seed, cs, 1000 + j, 100.0, 1e4)[None, :])
    for j, e in enumerate(idx["cl"]):
        params[els[e]["poff"]] = log_uniform(seed, cs, 5000 + j, 1e-12, 1e-9)
    _refresh(c, els, params)
    lane = np.stack([step, stop, step, 1e-12 * stop])
    return Batch(f"mos_chain_{p}", c, els, params, c.layout().n,
                 spec=dict(step=float(step[0]), stop=float(stop[0]), start=0.0, max_step=float(step[0])),
                 lane_spec=lane)


def mosfet_chains(B: int = 16384, patterns: int = 200, seed: int = 3, steps: int = 500) -> List[Batch]:
    """The uneven-size workload: B circuits over `patterns` distinct topologies."""
    per = [B // patterns + (1 if p < B % patterns else 0) for p in range(patterns)]
    return [mosfet_chain_batch(p, per[p], seed, steps) for p in range(patterns) if per[p] > 0]


# ---------------------------------------------------------------------------- config 5
def rlc_grid(w: int = 64, h: int = 64, draws: int = 16, seed: int = 5, source_at_far_corner: bool = False) -> Batch:
    """Complex AC sweep workload: a w x h grid mesh whose every edge is a
    series R-L branch (R log-U(0.1,10) ohm, L log-U(1 nH,1 uH); an internal
    node between them) and every node a capacitor to ground (C log-U(1 pF,
    1 nF)); a unit AC current source into the corner node (0,0); a 1 Mohm
    resistor from that node to ground gives the mesh its DC path (capacitors
    are not DC paths, circuit/circuit.go:717-719). `draws` value draws of the
    one topology; the transfer impedance is v at the opposite corner.
    source_at_far_corner drives the opposite corner instead (the same network
    and values: the reciprocity check)."""
    b = kc.Builder(f"rlc_grid_{w}x{h}")
    node = lambda x, y: f"g{x}_{y}"
    src = node(w - 1, h - 1) if source_at_far_corner else node(0, 0)
    b.isource("IAC", "0", src, 0.0, ac=1.0)
    b.resistor("RDC", node(0, 0), "0", 1e6)
    ridx, lidx, cidx = [], [], []
    for y in range(h):
        for x in range(w):
            for (x2, y2, tag) in ((x + 1, y, "h"), (x, y + 1, "v")):
                if x2 >= w or y2 >= h:
                    continue
                mid = f"m{tag}{x}_{y}"
                b.resistor(f"R{tag}{x}_{y}", node(x, y), mid, 1.0); ridx.append(len(b.elems) - 1)
                b.inductor(f"L{tag}{x}_{y}", mid, node(x2, y2), 1e-8); lidx.append(len(b.elems) - 1)
            b.capacitor(f"C{x}_{y}", node(x, y), "0", 1e-11); cidx.append(len(b.elems) - 1)
    c = b.build()
    els, params = _template(c, draws)
    cs = np.arange(draws)
    _set_resistors(els, params, ridx, np.stack([log_uniform(seed, cs, 3 * k, 0.1, 10.0) for k in range(len(ridx))]))
    for k, e in enumerate(lidx):
        params[els[e]["poff"]] = log_uniform(seed, cs, 3 * k + 1, 1e-9, 1e-6)
    for k, e in enumerate(cidx):
        params[els[e]["poff"]] = log_uniform(seed, cs, 3 * k + 2, 1e-12, 1e-9)
    _refresh(c, els, params)
    return Batch(f"rlc_grid_{w}x{h}_ac", c, els, params, c.layout().n)


def rlc_grid_frequencies(nf: int = 1024) -> np.ndarray:
    """nf log-spaced frequencies from 1 kHz to 10 GHz."""
    return np.logspace(3.0, 10.0, nf)


# ---------------------------------------------------------------------------- config 4
AMP_GENES = 60
AMP_LADDERS = (63, 63, 73)  # sections of the bias, inter-stage and load ladders (200 nodes in all)


def bjt_amplifier_circuit(seed: int = 4) -> Tuple[kc.Circuit, List[List[int]]]:
    """The learned-warm-start topology: a two-stage common-emitter BJT amplifier
    (BF U(50, 300), IS log-U(1e-16, 1e-14) A, drawn once from the seed) with a
    resistive ladder from the supply to the first base (bias), a resistive
    ladder from the first collector to the second base (DC coupling) and a
    resistive ladder as the second stage's load: 200 nodes. Returns the circuit
    and, per gene, the resistor elements it scales (genes 0-3: the stages' RC and
    RE; genes 4-59: groups of consecutive ladder sections)."""
    one = np.arange(1)
    bf = [lin_uniform(seed, one, 900 + k, 50.0, 300.0)[0] for k in range(2)]
    is_ = [log_uniform(seed, one, 910 + k, 1e-16, 1e-14)[0] for k in range(2)]
    b = kc.Builder("bjt two-stage amplifier")
    b.vsource("VCC", "vcc", "0", 12.0)
    groups: List[List[int]] = [[] for _ in range(AMP_GENES)]

    def res(name, a, c, r, gene):
        b.resistor(name, a, c, r)
        groups[gene].append(len(b.elems) - 1)

    def ladder(tag, start, end, sections, rs, rp, gene0, ngenes):
        prev = start
        for k in range(sections):
            node = end if (k == sections - 1 and end is not None) else f"{tag}{k + 1}"
            g = gene0 + (k * ngenes) // sections
            res(f"R{tag}s{k + 1}", prev, node, rs, g)
            res(f"R{tag}p{k + 1}", node, "0", rp, g)
            prev = node
        return prev

    n1, n2, n3 = AMP_LADDERS
    ladder("a", "vcc", "b1", n1, 300.0, 3.9e5, 4, 18)
    res("RC1", "vcc", "c1", 4.7e3, 0)
    res("RE1", "e1", "0", 1.0e3, 1)
    b.bjt("Q1", "c1", "b1", "e1", "qa")
    ladder("b", "c1", "b2", n2, 200.0, 1.0e6, 22, 19)
    res("RC2", "vcc", "c2", 2.2e3, 2)
    res("RE2", "e2", "0", 1.5e3, 3)
    b.bjt("Q2", "c2", "b2", "e2", "qb")
    ladder("c", "c2", None, n3, 100.0, 4.7e5, 41, 19)
    for k, name in enumerate(("qa", "qb")):
        # no series resistances: every unknown but the supply's branch current is a
        # node voltage, so a NodeSet guess determines the whole starting point
        b.model(name, "npn", {"is": is_[k], "bf": bf[k], "vaf": 80.0})
    return b.build(), groups


def bjt_amplifiers(B: int = 65536, seed: int = 4, first_circuit: int = 0) -> Tuple[Batch, np.ndarray]:
    """B draws of the amplifier's 60 genes, each log-uniform over one decade
    either side of its nominal value (the scale factor of its resistors).
    Returns the batch (DC OP) and the genes as log10 scale factors [60, B]
    in [-1, 1] (the network's input)."""
    c, groups = bjt_amplifier_circuit(seed)
    els, params = _template(c, B)
    cs = np.arange(first_circuit, first_circuit + B)
    genes = np.stack([lin_uniform(seed, cs, f, -1.0, 1.0) for f in range(AMP_GENES)])
    for g, idx in enumerate(groups):
        for k in idx:
            r = params[els[k]["poff"], 0] * 10.0 ** genes[g]
            _set_resistors(els, params, [k], r[None, :])
    return Batch("bjt_amplifier_dc_op", c, els, params, c.layout().n), genes


def circuit_of(batch: Batch, c: int) -> Tuple[np.ndarray, np.ndarray]:
    """Element table and parameter vector of circuit c of a batch."""
    return batch.elements, batch.circuit_params(c)


def cmos_latch() -> kc.Circuit:
    """Two cross-coupled level-1 CMOS inverters (a bistable latch) with 1 MΩ
    loads, 5 V supply. Three operating points: the metastable q = qb, and the
    two mirrored stable states; a cold start lands on the metastable one, so
    the starting guess (Options.NodeSet) decides which state Newton reaches."""
    nm = {"level": 1, "tox": 20e-9, "nsub": 3e16, "uo": 550, "lambda": 0.03}
    pm = {"level": 1, "tox": 20e-9, "nsub": 1e16, "uo": 200, "lambda": 0.05}
    return (kc.Builder("latch").vsource("VDD", "vdd", "0", 5)
            .mosfet("MN1", "qb", "q", "0", "0", "nm", {"l": 1e-6, "w": 10e-6})
            .mosfet("MP1", "qb", "q", "vdd", "vdd", "pm", {"l": 1e-6, "w": 25e-6})
            .mosfet("MN2", "q", "qb", "0", "0", "nm", {"l": 1e-6, "w": 10e-6})
            .mosfet("MP2", "q", "qb", "vdd", "vdd", "pm", {"l": 1e-6, "w": 25e-6})
            .resistor("R1", "q", "0", 1e6).resistor("R2", "qb", "0", 1e6)
            .model("nm", "nmos", nm).model("pm", "pmos", pm)).build()


def write_batch_file(path: str, batches: List[Batch], transient: bool = True) -> int:
    """Writes batches (one group each) in the KHB1 format read by kh_bench:
    "KHB1", int32 groups, int32 transient, double step, stop, max_step (of the
    first batch), then per group int32 num_nodes, nel, nparams, B, has_lane,
    the element table, params [P, B] and lane_spec [4, B] if present.
    Returns the number of circuits written."""
    import struct
    spec = batches[0].spec or {"step": 0.0, "stop": 0.0, "max_step": 0.0}
    total = 0
    with open(path, "wb") as f:
        f.write(b"KHB1")
        f.write(struct.pack("<ii3d", len(batches), 1 if transient else 0, spec["step"], spec["stop"],
                            spec["max_step"]))
        for b in batches:
            has_lane = transient and b.lane_spec is not None
            f.write(struct.pack("<5i", b.circuit.num_nodes, len(b.elements), b.params.shape[0], b.B,
                                1 if has_lane else 0))
            f.write(np.ascontiguousarray(b.elements).tobytes())
            f.write(np.ascontiguousarray(b.params, np.float64).tobytes())
            if has_lane:
                f.write(np.ascontiguousarray(b.lane_spec, np.float64).tobytes())
            total += b.B
    return total


def _main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="write a benchmark workload as a KHB1 batch file")
    ap.add_argument("workload", choices=["rectifier", "ladders", "mosfet_chains"])
    ap.add_argument("path")
    ap.add_argument("--batch", type=int, default=0, help="circuits (0: the workload's default)")
    ap.add_argument("--patterns", type=int, default=200)
    a = ap.parse_args(argv)
    if a.workload == "rectifier":
        bs, tran = [rectifier_rc_ladders(a.batch or 4096)], True
    elif a.workload == "ladders":
        bs, tran = [resistor_ladders(a.batch or 1024)], False
    else:
        bs, tran = mosfet_chains(a.batch or 16384, patterns=a.patterns), True
    print(write_batch_file(a.path, bs, tran), "circuits written to", a.path)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
