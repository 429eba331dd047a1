"""Circuit description and host-side device setup.

Mirrors kronspice's in-memory circuit (circuit/builder.go, circuit/circuit.go):
node indices are assigned in order of first appearance with ground = 0
(circuit/builder.go:38-48); elements keep their insertion order, which is the
device order of device.FromCircuit (device/device.go:347-451) and therefore the
MNA unknown numbering and the stamp order.

`flatten()` turns a circuit into the flat element table and parameter vector
of the C ABI (include/kronhip.h).  The parameter vector is the output of the
reference's setup step: model defaults (SPICE3), area factors, the diode's
derived constants (device/diode.go:85-128), MOSFET process-parameter
derivation (device/mosfet.go:128-323) and waveform defaults for the analysis
(device/waveform.go:25-66).  This is setup, run once per circuit; the solve path
itself (stamping, factorization, Newton, time stepping) runs on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Element kinds (include/kronhip.h)
R, C, L, V, I, D, M, Q = 1, 2, 3, 4, 5, 6, 7, 8
E, F, G, H = 10, 11, 12, 13  # VCVS, CCCS, VCCS, CCVS
K = 14  # mutual inductance
KIND_NAMES = {R: "R", C: "C", L: "L", V: "V", I: "I", D: "D", M: "M", Q: "Q", E: "E", F: "F", G: "G", H: "H",
              K: "K"}
NPAR = {R: 2, C: 3, L: 3, V: 12, I: 12, D: 16, M: 58, Q: 52, E: 1, F: 1, G: 1, H: 1, K: 1}

ELEMENT_DTYPE = np.dtype([("kind", "<i4"), ("node", "<i4", (4,)), ("poff", "<i4")])

# Physical constants, device/diode.go:14-19 and device/mosfet.go:116-120.
CHARGE_Q = 1.602176634e-19
BOLTZMANN_K = 1.380649e-23
TNOM_K = 300.15
NI_SI = 1.45e16


def _exact(expr: str) -> float:
    """A constant expression over the decimal constants above, evaluated
    exactly and rounded once, as Go folds untyped constant expressions."""
    from fractions import Fraction as F
    env = {"q": F("1.602176634e-19"), "k": F("1.380649e-23"), "T": F("300.15"), "F": F}
    return float(eval(expr, {"__builtins__": {}}, env))


EPS_OX = _exact("F('3.9') * F('8.854214871e-12')")
EPS_SI = _exact("F('11.7') * F('8.854214871e-12')")
VT_NOM = _exact("k * T / q")                          # BoltzmanK * TNomK / ChargeQ
EG_NOM = _exact("F('1.16') - F('7.02e-4') * T * T / (T + 1108)")
TWO_EPS_SI_Q = _exact("2 * F('11.7') * F('8.854214871e-12') * q")   # 2*epsSi*ChargeQ
TWO_EPS_SI = _exact("2 * F('11.7') * F('8.854214871e-12')")          # 2*epsSi
TNOM_R300 = _exact("T / 300")                                         # TNomK / 300
INV300_MINUS_INVT = _exact("F(1, 300) - 1 / T")                       # 1.0/300 - 1/TNomK
K_OVER_Q = _exact("k / q")                                            # BoltzmanK / ChargeQ
EPS_SI_Q = _exact("F('11.7') * F('8.854214871e-12') * q")           # epsSi * ChargeQ

WAVE_NONE, WAVE_PULSE, WAVE_SIN, WAVE_EXP = 0, 1, 2, 3


@dataclass
class Wave:
    """A source waveform: kind and the SPICE argument list as written."""
    kind: int
    args: Tuple[float, ...]


@dataclass
class Model:
    name: str
    type: str  # "d", "nmos", "pmos", "npn", "pnp"
    params: Dict[str, float]


@dataclass
class Element:
    name: str
    kind: int
    nodes: Tuple[int, ...]
    value: float = 0.0
    has_dc: bool = False
    wave: Optional[Wave] = None
    ac_mag: float = 0.0
    ac_phase: float = 0.0
    ic: float = 0.0
    has_ic: bool = False
    model: str = ""
    area: float = 0.0
    params: Dict[str, float] = field(default_factory=dict)
    ctrl_nodes: Tuple[int, ...] = ()  # E, G: nc+, nc- (node numbers)
    ctrl_source: str = ""             # F, H: the controlling element's name, lower-cased
    couples: Tuple[str, ...] = ()     # K: the two coupled inductors' names, lower-cased


class Circuit:
    """A built circuit: node names (index = node number), elements, models."""

    def __init__(self, title: str, names: List[str], elements: List[Element], models: Dict[str, Model]):
        self.title = title
        self.node_names = names
        self.elements = elements
        self.models = models

    @property
    def num_nodes(self) -> int:
        return len(self.node_names)

    def node(self, name: str) -> int:
        key = name.strip().lower()
        if key == "gnd":
            key = "0"
        return self.node_names.index(key)

    def element_index(self, name: str) -> int:
        for k, e in enumerate(self.elements):
            if e.name.lower() == name.lower():
                return k
        raise KeyError(name)

    # -- MNA layout ---------------------------------------------------------------
    def layout(self) -> "Layout":
        """Unknown numbering of device.SetupContext (device/device.go:36-100)."""
        nn = self.num_nodes
        branches = 0
        internal: List[int] = []
        branch_of: Dict[int, int] = {}
        internal_of: Dict[int, Tuple[int, ...]] = {}

        def alloc() -> int:
            nonlocal branches
            i = nn - 1 + branches
            branches += 1
            return i

        for k, e in enumerate(self.elements):
            if e.kind in (V, L, E, H):
                branch_of[k] = alloc()
            elif e.kind == D:
                p = diode_params(self, e)
                if p[14] > 0:  # grs
                    j = alloc()
                    internal.append(j)
                    internal_of[k] = (j,)
            elif e.kind == M:
                p = mosfet_params(self, e)
                extra = []
                if p[7] > 0:
                    j = alloc(); internal.append(j); extra.append(j)
                if p[8] > 0:
                    j = alloc(); internal.append(j); extra.append(j)
                if extra:
                    internal_of[k] = tuple(extra)
            elif e.kind == Q:  # cp, bp, ep behind RC, RB, RE (device/bjt.go:250-268)
                p = bjt_params(self, e)
                extra = []
                for g in (p[18], p[14], p[17]):
                    if g > 0:
                        j = alloc(); internal.append(j); extra.append(j)
                if extra:
                    internal_of[k] = tuple(extra)
        n = nn - 1 + branches
        is_v = np.zeros(n, dtype=bool)
        is_v[: nn - 1] = True
        for j in internal:
            is_v[j] = True
        return Layout(n=n, branch_of=branch_of, internal_of=internal_of, is_voltage=is_v)

    def flatten(self, step: float = 0.0, stop: float = 0.0) -> Tuple[np.ndarray, np.ndarray]:
        """Element table (ELEMENT_DTYPE) and parameter vector for one analysis.

        step/stop are the transient print step and stop time (0 outside a
        transient), used for SPICE waveform defaults (device/waveform.go:25-66).
        """
        els = np.zeros(len(self.elements), dtype=ELEMENT_DTYPE)
        params: List[float] = []
        for k, e in enumerate(self.elements):
            els[k]["kind"] = e.kind
            nodes = list(e.nodes) + [0] * (4 - len(e.nodes))  # a BJT without substrate: ground
            if e.kind in (E, G):
                nodes = [e.nodes[0], e.nodes[1], e.ctrl_nodes[0], e.ctrl_nodes[1]]
            elif e.kind in (F, H):  # node[2]: the controlling element's index (Link, device/controlled.go:67-77)
                try:
                    ctl = self.element_index(e.ctrl_source)
                except KeyError:
                    raise ValueError(f"{e.name}: controlling source {e.ctrl_source} not found") from None
                nodes = [e.nodes[0], e.nodes[1], ctl, 0]
            elif e.kind == K:  # node[0], node[1]: the coupled inductors' indices (Link, device/mutual.go:57-72)
                nodes = [self.element_index(n) for n in e.couples] + [0, 0]
            els[k]["node"] = nodes[:4]
            els[k]["poff"] = len(params)
            params.extend(element_params(self, e, step, stop))
        return els, np.asarray(params, dtype=np.float64)


@dataclass
class Layout:
    n: int
    branch_of: Dict[int, int]
    internal_of: Dict[int, Tuple[int, ...]]
    is_voltage: np.ndarray


class Builder:
    """Programmatic circuit construction, as circuit.Builder (circuit/builder.go:17-300)."""

    def __init__(self, title: str = ""):
        self.title = title
        self.index: Dict[str, int] = {"0": 0, "gnd": 0}
        self.names: List[str] = ["0"]
        self.elems: List[Element] = []
        self.models: Dict[str, Model] = {}

    def node(self, name: str) -> int:
        key = name.strip().lower()
        if key in self.index:
            return self.index[key]
        i = len(self.names)
        self.index[key] = i
        self.names.append(key)
        return i

    def add(self, e: Element, *names: str) -> "Builder":
        e.nodes = tuple(self.node(n) for n in names)
        self.elems.append(e)
        return self

    def model(self, name: str, typ: str, params: Optional[Dict[str, float]] = None) -> "Builder":
        self.models[name.lower()] = Model(name.lower(), typ.lower(), {k.lower(): float(v) for k, v in (params or {}).items()})
        return self

    def resistor(self, name, a, b, r):
        return self.add(Element(name, R, (), value=float(r)), a, b)

    def capacitor(self, name, a, b, c, ic: Optional[float] = None):
        return self.add(Element(name, C, (), value=float(c), ic=ic or 0.0, has_ic=ic is not None), a, b)

    def inductor(self, name, a, b, l, ic: Optional[float] = None):
        return self.add(Element(name, L, (), value=float(l), ic=ic or 0.0, has_ic=ic is not None), a, b)

    def vsource(self, name, a, b, dc=0.0, wave: Optional[Wave] = None, ac: float = 0.0, ac_phase: float = 0.0,
                has_dc: Optional[bool] = None):
        hd = (wave is None) if has_dc is None else has_dc
        return self.add(Element(name, V, (), value=float(dc), has_dc=hd, wave=wave, ac_mag=ac, ac_phase=ac_phase), a, b)

    def isource(self, name, a, b, dc=0.0, wave: Optional[Wave] = None, ac: float = 0.0, ac_phase: float = 0.0,
                has_dc: Optional[bool] = None):
        hd = (wave is None) if has_dc is None else has_dc
        return self.add(Element(name, I, (), value=float(dc), has_dc=hd, wave=wave, ac_mag=ac, ac_phase=ac_phase), a, b)

    def diode(self, name, anode, cathode, model, area: float = 0.0):
        return self.add(Element(name, D, (), model=model.lower(), area=area), anode, cathode)

    def mosfet(self, name, drain, gate, source, bulk, model, params: Optional[Dict[str, float]] = None):
        # AddMulti(e, drain, source, gate, bulk): circuit/builder.go:113-118
        e = Element(name, M, (), model=model.lower(), params={k.lower(): float(v) for k, v in (params or {}).items()})
        return self.add(e, drain, source, gate, bulk)

    def controlled(self, name, kind, p, n, gain, ctrl_nodes=None, ctrl_source: str = ""):
        """E and G (ctrl_nodes = (nc+, nc-)), F and H (ctrl_source): the control
        nodes are numbered before the output nodes, as circuit.FromDeck and
        Builder.VCVS / VCCS register them (circuit/builder.go:119-127, :409-412)."""
        e = Element(name, kind, (), value=float(gain), ctrl_source=ctrl_source.lower())
        if kind in (E, G):
            e.ctrl_nodes = (self.node(ctrl_nodes[0]), self.node(ctrl_nodes[1]))
        return self.add(e, p, n)

    def vcvs(self, name, p, n, cp, cn, gain):
        return self.controlled(name, E, p, n, gain, ctrl_nodes=(cp, cn))

    def vccs(self, name, p, n, cp, cn, gm):
        return self.controlled(name, G, p, n, gm, ctrl_nodes=(cp, cn))

    def cccs(self, name, p, n, vname, gain):
        return self.controlled(name, F, p, n, gain, ctrl_source=vname)

    def ccvs(self, name, p, n, vname, r):
        return self.controlled(name, H, p, n, r, ctrl_source=vname)

    def mutual(self, name, l1, l2, k):
        """Mutual inductance (SPICE K) between inductors l1 and l2, coupling
        coefficient k; its nodes are ground (circuit/builder.go:80-84)."""
        e = Element(name, K, (), value=float(k), couples=(l1.lower(), l2.lower()))
        return self.add(e, "0", "0")

    def bjt(self, name, collector, base, emitter, model, substrate: Optional[str] = None, area: float = 0.0):
        # AddMulti(e, collector, emitter, base[, substrate]): circuit/builder.go:99-103, netlist
        # element 'Q' (circuit/builder.go:458-461)
        e = Element(name, Q, (), model=model.lower(), area=area)
        names = (collector, emitter, base) + ((substrate,) if substrate is not None else ())
        return self.add(e, *names)

    def build(self) -> Circuit:
        return Circuit(self.title, list(self.names), list(self.elems), dict(self.models))


# --------------------------------------------------------------------------------
# Setup: element -> resolved parameter block (layouts of include/kronhip.h)
# --------------------------------------------------------------------------------

def resolve_wave(w: Optional[Wave], step: float, stop: float) -> Tuple[int, List[float]]:
    """resolveWave, device/waveform.go:25-66."""
    if w is None:
        return WAVE_NONE, [0.0] * 7
    if step <= 0:
        step = 1e-9
    if stop <= 0:
        stop = 1e-6
    args = list(w.args)

    def arg(i, d):
        return args[i] if i < len(args) else d

    if w.kind == WAVE_PULSE:
        a = [arg(0, 0.0), arg(1, 0.0), arg(2, 0.0), arg(3, step), arg(4, step), arg(5, stop), arg(6, stop)]
        if a[3] == 0:
            a[3] = step
        if a[4] == 0:
            a[4] = step
        return WAVE_PULSE, a
    if w.kind == WAVE_SIN:
        return WAVE_SIN, [arg(0, 0.0), arg(1, 0.0), arg(2, 1 / stop), arg(3, 0.0), arg(4, 0.0), arg(5, 0.0), 0.0]
    if w.kind == WAVE_EXP:
        td1 = arg(2, 0.0)
        a = [arg(0, 0.0), arg(1, 0.0), td1, arg(3, step), arg(4, td1 + step), arg(5, step), 0.0]
        if a[3] == 0:
            a[3] = step
        if a[5] == 0:
            a[5] = step
        return WAVE_EXP, a
    raise ValueError(f"unsupported waveform kind {w.kind}")


def _model(c: Circuit, e: Element) -> Model:
    m = c.models.get(e.model)
    if m is None:
        raise ValueError(f"{e.name} has no model")
    return m


def breakdown_voltage(is_: float, bv: float, ibv: float, vt: float) -> float:
    """breakdownVoltage, device/diode.go:166-187 (bisection)."""
    if ibv < is_ * bv / vt:
        return bv

    def g(x):
        return is_ * (math.exp(min((bv - x) / vt, 700)) - 1 + x / vt) - ibv

    lo, hi = min(0.0, bv - vt * math.log1p(ibv / is_)) - 10 * vt, bv
    i = 0
    while i < 200 and g(lo) <= 0:
        lo -= 10 * vt
        i += 1
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if g(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def diode_params(c: Circuit, e: Element) -> List[float]:
    """NewDiode, device/diode.go:85-128 -> the D layout of include/kronhip.h."""
    m = _model(c, e).params
    p = dict(IS=1e-14, N=1.0, RS=0.0, CJO=0.0, VJ=1.0, M=0.5, TT=0.0, BV=0.0, IBV=1e-3, FC=0.5)
    for key, name in (("is", "IS"), ("n", "N"), ("rs", "RS"), ("cjo", "CJO"), ("cj0", "CJO"), ("vj", "VJ"),
                      ("m", "M"), ("tt", "TT"), ("ibv", "IBV"), ("fc", "FC")):
        if key in m:
            p[name] = m[key]
    has_bv = "bv" in m
    if has_bv:
        p["BV"] = m["bv"]
    area = e.area if e.area != 0 else 1.0
    p["IS"] *= area
    p["CJO"] *= area
    p["IBV"] *= area
    if p["RS"] > 0:
        p["RS"] /= area
    vte = p["N"] * BOLTZMANN_K * TNOM_K / CHARGE_Q
    vcrit = vte * math.log(vte / (math.sqrt(2) * p["IS"]))
    grs = 1 / p["RS"] if p["RS"] > 0 else 0.0
    has_charge = p["CJO"] > 0 or p["TT"] > 0
    xbv = breakdown_voltage(p["IS"], p["BV"], p["IBV"], BOLTZMANN_K * TNOM_K / CHARGE_Q) if has_bv else 0.0
    return [p["IS"], p["N"], p["RS"], p["CJO"], p["VJ"], p["M"], p["TT"], p["BV"], p["IBV"], p["FC"],
            1.0 if has_bv else 0.0, vte, vcrit, xbv, grs, 1.0 if has_charge else 0.0]


def depl_f(vj: float, m: float, fc: float) -> Tuple[float, float, float]:
    """deplF, device/bjt.go:245-247."""
    return vj * (1 - (1 - fc) ** (1 - m)) / (1 - m), (1 - fc) ** (1 + m), 1 - fc * (1 + m)


def mosfet_params(c: Circuit, e: Element) -> List[float]:
    """NewMOSFET + derive for level 1 (device/mosfet.go:128-323) -> the M layout."""
    mod = _model(c, e)
    mp = mod.params
    level = int(mp.get("level", 1))
    if level not in (1, 2, 3):
        raise ValueError(f"{e.name}: MOSFET level {level} is not implemented")
    typ = -1.0 if mod.type == "pmos" else 1.0
    p = dict(VTO=0.0, KP=2e-5, GAMMA=0.0, PHI=0.6, LAMBDA=0.0, RD=0.0, RS=0.0, CBD=0.0, CBS=0.0, IS=1e-14, PB=0.8,
             CGSO=0.0, CGDO=0.0, CGBO=0.0, RSH=0.0, CJ=0.0, MJ=0.5, CJSW=0.0, MJSW=0.5, JS=0.0, TOX=0.0, NSUB=0.0,
             NSS=0.0, TPG=1.0, LD=0.0, UO=600.0, NFS=0.0, XJ=0.0, UCRIT=1e4, UEXP=0.0, UTRA=0.0, VMAX=0.0, NEFF=1.0,
             DELTA=0.0, THETA=0.0, ETA=0.0, KAPPA=0.0, WD=0.0)
    given = set()
    if level >= 2:  # mos2set, mos3set: MJSW 0.33, TOX 1e-7 (given)
        p["MJSW"] = 0.33
        p["TOX"] = 1e-7
        given.add("tox")
    keys = [("VTO", ("vto", "vt0")), ("KP", ("kp",)), ("GAMMA", ("gamma",)), ("PHI", ("phi",)),
            ("LAMBDA", ("lambda",)), ("RD", ("rd",)), ("RS", ("rs",)), ("CBD", ("cbd",)), ("CBS", ("cbs",)),
            ("IS", ("is",)), ("PB", ("pb",)), ("CGSO", ("cgso",)), ("CGDO", ("cgdo",)), ("CGBO", ("cgbo",)),
            ("RSH", ("rsh",)), ("CJ", ("cj",)), ("MJ", ("mj",)), ("CJSW", ("cjsw",)), ("MJSW", ("mjsw",)),
            ("JS", ("js",)), ("TOX", ("tox",)), ("NSUB", ("nsub",)), ("NSS", ("nss",)), ("NFS", ("nfs",)),
            ("TPG", ("tpg",)), ("XJ", ("xj",)), ("LD", ("ld",)), ("UO", ("uo", "u0")), ("UCRIT", ("ucrit",)),
            ("UEXP", ("uexp",)), ("UTRA", ("utra",)), ("VMAX", ("vmax",)), ("NEFF", ("neff",)), ("FC", ("fc",)),
            ("DELTA", ("delta",)), ("THETA", ("theta",)), ("ETA", ("eta",)), ("KAPPA", ("kappa",)), ("WD", ("wd",))]
    p["FC"] = 0.5
    for name, ks in keys:
        for k in ks:
            if k in mp:
                p[name] = mp[k]
                given.add(ks[0])
                break
    if level == 3 and "kappa" not in given:
        p["KAPPA"] = 0.2
    inst = dict(l=100e-6, w=100e-6, ad=0.0, as_=0.0, pd=0.0, ps=0.0, nrd=0.0, nrs=0.0, m=1.0)
    for k, v in e.params.items():
        kk = "as_" if k == "as" else k
        if kk in inst:
            inst[kk] = v
    vt = VT_NOM
    leff = inst["l"] - 2 * p["LD"]
    weff = inst["w"] - 2 * p["WD"] if level == 3 else inst["w"]
    if not (leff > 0) or not (weff > 0):
        raise ValueError(f"{e.name}: effective channel length or width is not positive")
    cox = 0.0
    if "tox" in given and p["TOX"] > 0:
        cox = EPS_OX / p["TOX"]
        if "kp" not in given:
            p["KP"] = p["UO"] * cox * 1e-4
        if "nsub" in given:
            nsub = p["NSUB"] * 1e6
            ni = NI_SI
            if level == 3:  # mos3temp: ni scaled from 300 K to TNOM
                eg = EG_NOM
                r = TNOM_R300
                ni *= r * math.sqrt(r) * math.exp(0.5 * eg * INV300_MINUS_INVT / K_OVER_Q)
            if not (nsub > ni):
                raise ValueError("NSUB below the intrinsic density")
            if "phi" not in given:
                p["PHI"] = max(0.1, 2 * vt * math.log(nsub / ni))
            if "gamma" not in given:
                p["GAMMA"] = math.sqrt(TWO_EPS_SI_Q * nsub) / cox
            if "vto" not in given:
                eg = EG_NOM
                fermis = typ * 0.5 * p["PHI"]
                wkfng = 3.2
                if p["TPG"] != 0:
                    fermig = typ * p["TPG"] * 0.5 * eg
                    wkfng = 3.25 + 0.5 * eg - fermig
                wkfngs = wkfng - (3.25 + 0.5 * eg + fermis)
                vfb = wkfngs - p["NSS"] * 1e4 * CHARGE_Q / cox
                p["VTO"] = vfb + typ * (p["GAMMA"] * math.sqrt(p["PHI"]) + p["PHI"])
            if "cj" not in given:
                p["CJ"] = math.sqrt(EPS_SI_Q * nsub / (2 * p["PB"]))
    mult = inst["m"]
    oxide_cap = cox * leff * weff * mult
    beta = p["KP"] * weff / leff * mult
    vto = typ * p["VTO"]
    vbi = vto - p["GAMMA"] * math.sqrt(p["PHI"])
    rd, rs = p["RD"], p["RS"]
    if "rd" not in given and p["RSH"] > 0:
        rd = p["RSH"] * inst["nrd"]
    if "rs" not in given and p["RSH"] > 0:
        rs = p["RSH"] * inst["nrs"]
    gdpr = mult / rd if rd > 0 else 0.0
    gspr = mult / rs if rs > 0 else 0.0
    isd = iss = p["IS"] * mult
    if p["JS"] > 0 and inst["ad"] > 0 and inst["as_"] > 0:
        isd, iss = p["JS"] * inst["ad"] * mult, p["JS"] * inst["as_"] * mult
    vcrit_d = vt * math.log(vt / (math.sqrt(2) * max(isd, 1e-300)))
    vcrit_s = vt * math.log(vt / (math.sqrt(2) * max(iss, 1e-300)))
    czbd = p["CBD"] * mult if "cbd" in given else p["CJ"] * inst["ad"] * mult
    czbs = p["CBS"] * mult if "cbs" in given else p["CJ"] * inst["as_"] * mult
    czbdsw = p["CJSW"] * inst["pd"] * mult
    czbssw = p["CJSW"] * inst["ps"] * mult
    fc = p["FC"]
    fbt = depl_f(p["PB"], p["MJ"], fc)
    fsw = depl_f(p["PB"], p["MJSW"], fc)
    cgso, cgdo, cgbo = p["CGSO"] * weff * mult, p["CGDO"] * weff * mult, p["CGBO"] * leff * mult
    has_j = czbd + czbdsw + czbs + czbssw > 0
    has_g = oxide_cap > 0 or cgso + cgdo + cgbo > 0
    # deriveLevel, device/mosfet_levels.go:22-55
    xd = alpha = cfs = factor = narrow = eta3 = sqrt_pb = 0.0
    has_nsub = "nsub" in given
    if has_nsub:
        alpha = TWO_EPS_SI / (CHARGE_Q * p["NSUB"] * 1e6)
        xd = math.sqrt(alpha)
    if level == 2:
        cfs = CHARGE_Q * p["NFS"] * 1e4 / cox
        factor = p["DELTA"] * math.pi * EPS_SI / (4 * cox * weff)
        sqrt_pb = math.sqrt(p["PB"])
        if p["VMAX"] > 0 and not (p["NEFF"] > 0):
            raise ValueError(f"{e.name}: NEFF must be positive")
    elif level == 3:
        cfs = CHARGE_Q * p["NFS"] * 1e4 / cox
        narrow = p["DELTA"] * 0.5 * math.pi * EPS_SI / cox
        eta3 = p["ETA"] * 8.15e-22 / (cox * leff * leff * leff)
    return [typ, beta, vbi, vto, p["GAMMA"], p["PHI"], p["LAMBDA"], gdpr, gspr, isd, iss, vcrit_d, vcrit_s,
            czbd, czbdsw, czbs, czbssw, p["PB"], p["MJ"], p["MJSW"], fc, fbt[0], fbt[1], fbt[2], fsw[0], fsw[1],
            fsw[2], cgso, cgdo, cgbo, oxide_cap, vt, 1.0 if has_j else 0.0, 1.0 if has_g else 0.0,
            float(level), p["NFS"], p["XJ"], p["UCRIT"], p["UEXP"], p["VMAX"], p["NEFF"], p["DELTA"], p["THETA"],
            p["ETA"], p["KAPPA"], p["UO"], leff, weff, cox, xd, alpha, cfs, factor, narrow, eta3, sqrt_pb,
            1.0 if has_nsub else 0.0, p["LD"]]


BJT_DEFAULTS = dict(IS=1e-16, BF=100.0, NF=1.0, VAF=0.0, IKF=0.0, ISE=0.0, NE=1.5, BR=1.0, NR=1.0, VAR=0.0,
                    IKR=0.0, ISC=0.0, NC=2.0, RB=0.0, IRB=0.0, RBM=0.0, RE=0.0, RC=0.0, CJE=0.0, VJE=0.75,
                    MJE=0.33, TF=0.0, XTF=0.0, VTF=0.0, ITF=0.0, CJC=0.0, VJC=0.75, MJC=0.33, XCJC=1.0, TR=0.0,
                    CJS=0.0, VJS=0.75, MJS=0.0, FC=0.5, ISS=0.0, NS=1.0, SUBS=0.0)
BJT_KEYS = (("IS", ("is",)), ("BF", ("bf",)), ("NF", ("nf",)), ("VAF", ("vaf", "va")), ("IKF", ("ikf", "ik")),
            ("ISE", ("ise",)), ("NE", ("ne",)), ("BR", ("br",)), ("NR", ("nr",)), ("VAR", ("var", "vb")),
            ("IKR", ("ikr",)), ("ISC", ("isc",)), ("NC", ("nc",)), ("RB", ("rb",)), ("IRB", ("irb",)))
BJT_KEYS2 = (("RE", ("re",)), ("RC", ("rc",)), ("CJE", ("cje",)), ("VJE", ("vje", "pe")), ("MJE", ("mje", "me")),
             ("TF", ("tf",)), ("XTF", ("xtf",)), ("VTF", ("vtf",)), ("ITF", ("itf",)), ("CJC", ("cjc",)),
             ("VJC", ("vjc", "pc")), ("MJC", ("mjc", "mc")), ("XCJC", ("xcjc",)), ("TR", ("tr",)),
             ("CJS", ("cjs", "ccs")), ("VJS", ("vjs", "ps")), ("MJS", ("mjs", "ms")), ("FC", ("fc",)),
             ("ISS", ("iss",)), ("NS", ("ns",)), ("SUBS", ("subs",)))
BJT_LAYOUT = ("IS", "BF", "NF", "VAF", "IKF", "ISE", "NE", "BR", "NR", "VAR", "IKR", "ISC", "NC", "RB", "IRB", "RBM",
              "RE", "RC", "CJE", "VJE", "MJE", "TF", "XTF", "VTF", "ITF", "CJC", "VJC", "MJC", "XCJC", "TR", "CJS",
              "VJS", "MJS", "FC", "ISS", "NS", "SUBS")


def bjt_params(c: Circuit, e: Element) -> List[float]:
    """NewBJT (device/bjt.go:142-242) -> the Q layout of include/kronhip.h."""
    mod = _model(c, e)
    mp = mod.params
    if mod.type not in ("npn", "pnp"):
        raise ValueError(f"{e.name} needs an NPN or PNP model, {mod.name!r} is type {mod.type}")
    p = dict(BJT_DEFAULTS)
    typ = -1.0 if mod.type == "pnp" else 1.0

    def get(keys):
        for name, ks in keys:
            for k in ks:
                if k in mp:
                    p[name] = mp[k]
                    break

    get(BJT_KEYS)
    p["RBM"] = p["RB"]
    get((("RBM", ("rbm",)),))
    get(BJT_KEYS2)
    if p["SUBS"] == 0:  # vertical NPN, lateral PNP
        p["SUBS"] = -1.0 if typ < 0 else 1.0
    if p["RBM"] > p["RB"]:
        p["RBM"] = p["RB"]
    area = e.area if e.area != 0 else 1.0
    for k in ("IS", "ISE", "ISC", "IKF", "IKR", "IRB", "ITF", "CJE", "CJC", "CJS", "ISS"):
        p[k] *= area
    for k in ("RB", "RBM", "RE", "RC"):
        p[k] /= area
    vt = VT_NOM
    vcrit = vt * math.log(vt / (math.sqrt(2) * p["IS"]))
    rbpr, rbpi = p["RBM"], p["RB"] - p["RBM"]
    gcpr = 1 / p["RC"] if p["RC"] > 0 else 0.0
    gepr = 1 / p["RE"] if p["RE"] > 0 else 0.0
    ovtf = 1 / (1.44 * p["VTF"]) if p["VTF"] > 0 else 0.0
    fe = depl_f(p["VJE"], p["MJE"], p["FC"])
    fcc = depl_f(p["VJC"], p["MJC"], p["FC"])
    has_charge = p["CJE"] > 0 or p["CJC"] > 0 or p["CJS"] > 0 or p["TF"] > 0 or p["TR"] > 0
    return ([typ] + [p[k] for k in BJT_LAYOUT] + [vt, vcrit, rbpr, rbpi, gcpr, gepr, ovtf] + list(fe) + list(fcc)
            + [1.0 if has_charge else 0.0])


def source_params(e: Element, step: float, stop: float) -> List[float]:
    """sourceValue (device/source.go:21-66) with the waveform resolved."""
    kind, a = resolve_wave(e.wave, step, stop)
    return [e.value, 1.0 if e.has_dc else 0.0, float(kind)] + list(a) + [e.ac_mag, e.ac_phase]


def element_params(c: Circuit, e: Element, step: float, stop: float) -> List[float]:
    if e.kind == R:
        return [e.value, 1 / e.value]
    if e.kind in (C, L):
        return [e.value, e.ic, 1.0 if e.has_ic else 0.0]
    if e.kind in (V, I):
        return source_params(e, step, stop)
    if e.kind == D:
        return diode_params(c, e)
    if e.kind == M:
        return mosfet_params(c, e)
    if e.kind == Q:
        return bjt_params(c, e)
    if e.kind in (E, F, G, H):
        return [e.value]  # the gain (device/controlled.go:45-51)
    if e.kind == K:
        return [mutual_inductance(c, e)]
    raise NotImplementedError(f"element kind {e.kind} ({e.name}) is not on this path")


def mutual_inductance(c: Circuit, e: Element) -> float:
    """M = k·sqrt(L1·L2) of a K element (device/mutual.go:44-50 with the inductances
    device/device.go:407-416 reads), after circuit.validateMutual's checks
    (circuit/circuit.go:804-830): 0 < k <= 1, two distinct linear inductors, one K
    per pair."""
    if not (e.value > 0 and e.value <= 1):
        raise ValueError(f"{e.name}: coupling coefficient {e.value:g} must be in (0, 1]")
    a, b = e.couples
    if a == b:
        raise ValueError(f"{e.name} couples {a} with itself")
    ls = []
    for n in (a, b):
        try:
            el = c.elements[c.element_index(n)]
        except KeyError:
            el = None
        if el is None or el.kind != L:
            raise ValueError(f"{e.name}: {n!r} is not a linear inductor")
        ls.append(el.value)
    pair = tuple(sorted((a, b)))
    for o in c.elements:
        if o is not e and o.kind == K and tuple(sorted(o.couples)) == pair and c.elements.index(o) < c.elements.index(e):
            raise ValueError(f"{e.name}: {pair[0]} and {pair[1]} are already coupled")
    return e.value * math.sqrt(ls[0] * ls[1])


def unknown_of_node(node: int) -> int:
    """SetupContext.Unknown: node k -> k - 1 (ground -> -1)."""
    return node - 1
