"""SPICE netlist subset reader for the solve path's element set.

Follows kronspice's netlist package for the cards this path runs
(netlist/parser.go, netlist/units.go): R, C, L, V, I, D, M and Q elements,
.model, .op, .tran, .ac, .dc and .end; comments (* lines and ; tails) and +
continuations; .param, .global, .subckt/.ends with X instances and
{expressions} through netlist_pre (netlist/preprocess.go, netlist/expr.go).  Elements are handed to the circuit builder in file order, with
multi-terminal node order as circuit.FromDeck (circuit/builder.go:307-470):
MOSFET drain, source, gate, bulk; BJT collector, emitter, base[, substrate].
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from . import circuit as kc
from .circuit import Builder, Circuit, Wave, WAVE_EXP, WAVE_PULSE, WAVE_SIN

_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def parse_value(s: str) -> float:
    """ParseValue, netlist/units.go:24-80: engineering suffix applied as a
    decimal exponent before conversion (one rounding)."""
    t = s.strip()
    m = _NUM.match(t)
    if not m:
        raise ValueError(f"bad value {s!r}")
    lit, rest = t[: m.end()], t[m.end():].lower()
    exp10, scale = 0, 1.0
    if rest.startswith("meg"):
        exp10 = 6
    elif rest.startswith("mil"):
        scale = 25.4e-6
    elif rest:
        exp10 = {"t": 12, "g": 9, "k": 3, "m": -3, "u": -6, "n": -9, "p": -12, "f": -15}.get(rest[0], 0)
        if exp10 == 0 and not rest[0].isalpha():
            raise ValueError(f"bad value {s!r}")
    if exp10:
        mant, e = lit, 0
        k = max(lit.find("e"), lit.find("E"))
        if k >= 0:
            mant, e = lit[:k], int(lit[k + 1:])
        lit = f"{mant}e{e + exp10}"
    v = float(lit) * scale
    return v


@dataclass
class Analysis:
    kind: str  # "op", "tran", "ac", "dc"
    args: Dict[str, object] = field(default_factory=dict)


@dataclass
class Deck:
    title: str
    circuit: Circuit
    analyses: List[Analysis]


_WAVES = {"pulse": WAVE_PULSE, "sin": WAVE_SIN, "exp": WAVE_EXP}


def _logical_lines(text: str) -> Tuple[str, List[str]]:
    lines = text.splitlines()
    title = lines[0].strip() if lines else ""
    out: List[str] = []
    for raw in lines[1:]:
        s = raw.rstrip("\r")
        t = s.lstrip(" \t")
        if t.startswith("+") and out:
            out[-1] += " " + t[1:]
            continue
        if not t or t.startswith("*"):
            continue
        out.append(t)
    cleaned = []
    for ln in out:
        k = ln.find(";")
        if k >= 0:
            ln = ln[:k]
        if ln.strip():
            cleaned.append(ln.strip())
    return title, cleaned


def _tokens(line: str) -> List[str]:
    line = re.sub(r"\s*=\s*", "=", line)
    return line.replace("(", " ").replace(")", " ").replace(",", " ").split()


def _source_spec(f: List[str]):
    """parseSourceSpec, netlist/parser.go:856-936."""
    value, has_dc, wave, ac_mag, ac_ph = 0.0, False, None, 0.0, 0.0
    i = 0

    def is_kw(tok):
        lo = tok.lower()
        return lo in ("dc", "ac") or lo in _WAVES or lo == "pwl"

    while i < len(f):
        lo = f[i].lower()
        if lo == "dc":
            i += 1
            value, has_dc = parse_value(f[i]), True
            i += 1
        elif lo == "ac":
            i += 1
            ac_mag = 1.0
            if i < len(f) and not is_kw(f[i]):
                ac_mag = parse_value(f[i]); i += 1
                if i < len(f) and not is_kw(f[i]):
                    ac_ph = parse_value(f[i]); i += 1
        elif lo in _WAVES:
            i += 1
            args = []
            while i < len(f) and not is_kw(f[i]):
                args.append(parse_value(f[i])); i += 1
            wave = Wave(_WAVES[lo], tuple(args))
        elif lo == "pwl":
            raise NotImplementedError("PWL sources are not on this path")
        else:
            value, has_dc = parse_value(f[i]), True
            i += 1
    return value, has_dc, wave, ac_mag, ac_ph


def parse(text: str) -> Deck:
    from .netlist_pre import expand
    title, lines = _logical_lines(text)
    # .param / .global / .subckt and X instances, {expressions}: flat lines
    # (netlist/parser.go:156-182, netlist/preprocess.go)
    lines, _params = expand(lines)
    b = Builder(title)
    analyses: List[Analysis] = []
    elements: List[List[str]] = []
    for ln in lines:
        tok = _tokens(ln)
        head = tok[0].lower()
        if head == ".end":
            break
        if head == ".model":
            name, typ = tok[1], tok[2].lower()
            params = {}
            for t in tok[3:]:
                if "=" in t:
                    k, v = t.split("=", 1)
                    params[k.lower()] = parse_value(v)
            b.model(name, typ, params)
        elif head == ".op":
            analyses.append(Analysis("op"))
        elif head == ".tran":
            args = [a for a in tok[1:] if a.lower() != "uic"]
            vals = [parse_value(a) for a in args]
            analyses.append(Analysis("tran", dict(step=vals[0], stop=vals[1],
                                                  start=vals[2] if len(vals) > 2 else 0.0,
                                                  max_step=vals[3] if len(vals) > 3 else 0.0,
                                                  uic=any(a.lower() == "uic" for a in tok[1:]))))
        elif head == ".ac":
            analyses.append(Analysis("ac", dict(variation=tok[1].lower(), points=int(parse_value(tok[2])),
                                                fstart=parse_value(tok[3]), fstop=parse_value(tok[4]))))
        elif head == ".dc":
            # source: the lower-cased name for lookups; source_name: as written, as
            # DCSweep.Source keeps it (netlist/parser.go:367-368)
            analyses.append(Analysis("dc", dict(source=tok[1].lower(), source_name=tok[1], start=parse_value(tok[2]),
                                                stop=parse_value(tok[3]), step=parse_value(tok[4]))))
        elif head.startswith("."):
            raise NotImplementedError(f"unsupported control card {head}")
        else:
            elements.append(tok)
    for tok in elements:
        name, k = tok[0], tok[0][0].upper()
        if k == "R":
            b.resistor(name, tok[1], tok[2], parse_value(tok[3]))
        elif k in ("C", "L"):
            ic = None
            val = parse_value(tok[3])
            for t in tok[4:]:
                if t.lower().startswith("ic="):
                    ic = parse_value(t[3:])
            (b.capacitor if k == "C" else b.inductor)(name, tok[1], tok[2], val, ic=ic)
        elif k in ("V", "I"):
            value, has_dc, wave, ac, ph = _source_spec(tok[3:])
            (b.vsource if k == "V" else b.isource)(name, tok[1], tok[2], value, wave=wave, ac=ac, ac_phase=ph,
                                                   has_dc=has_dc)
        elif k == "D":
            area = parse_value(tok[4]) if len(tok) > 4 and "=" not in tok[4] else 0.0
            b.diode(name, tok[1], tok[2], tok[3], area=area)
        elif k == "M":
            params = {}
            for t in tok[6:]:
                if "=" in t:
                    kk, v = t.split("=", 1)
                    params[kk.lower()] = parse_value(v)
            b.mosfet(name, tok[1], tok[2], tok[3], tok[4], tok[5], params)
        elif k == "Q":  # Qname c b e [s] model [area], netlist/parser.go:661-695
            t = tok[1:]
            if not 4 <= len(t) <= 6:
                raise ValueError(f"{name}: BJT needs: Qname c b e [s] model [area]")
            area, sub = 0.0, None
            if len(t) == 4:
                model = t[3]
            elif len(t) == 5:
                try:
                    area = parse_value(t[4])
                    model = t[3]
                except ValueError:
                    sub, model = t[3], t[4]
            else:
                sub, model, area = t[3], t[4], parse_value(t[5])
            b.bjt(name, t[0], t[1], t[2], model, substrate=sub, area=area)
        elif k in ("E", "G"):  # Ename n+ n- nc+ nc- gain, netlist/parser.go:468-478
            if len(tok) != 6:
                raise ValueError(f"{name}: {k} element needs: name n+ n- nc+ nc- gain")
            b.controlled(name, kc.E if k == "E" else kc.G, tok[1], tok[2], parse_value(tok[5]),
                         ctrl_nodes=(tok[3], tok[4]))
        elif k in ("F", "H"):  # Fname n+ n- vname gain, netlist/parser.go:620-630
            if len(tok) != 5:
                raise ValueError(f"{name}: {k} element needs: name n+ n- vname gain")
            b.controlled(name, kc.F if k == "F" else kc.H, tok[1], tok[2], parse_value(tok[4]), ctrl_source=tok[3])
        elif k == "K":  # Kname L1 L2 k, netlist/parser.go:589-621 (no spin-connection terms here)
            if len(tok) != 4:
                raise ValueError(f"{name}: K element needs: Kname L1 L2 coupling")
            b.mutual(name, tok[1], tok[2], parse_value(tok[3]))
        else:
            raise NotImplementedError(f"element {name} is not on this path")
    return Deck(title, b.build(), analyses)


def ac_frequencies(variation: str, points: int, fstart: float, fstop: float) -> List[float]:
    """ACSpec.Frequencies, analysis/ac.go:32-69."""
    import math
    if variation == "lin":
        if points == 1:
            return [fstart]
        step = (fstop - fstart) / (points - 1)
        return [fstart + k * step for k in range(points)]
    base = 10.0 if variation == "dec" else 2.0
    ratio = base ** (1 / points)
    n = int(math.floor(math.log(fstop / fstart) / math.log(ratio) + 1e-9)) + 1
    return [fstart * ratio ** k for k in range(n)]
