"""Netlist preprocessing: .param, .global and .subckt/.ends with X instances,
expressions in {braces} and 'quotes' (netlist/preprocess.go, netlist/expr.go).

The result is the flat list of element and control lines the element parser
reads, as the reference's expander produces it (netlist/preprocess.go:
412-720): parameters are scoped lexically (an instance sees its .subckt
defaults, the values on its X line, the .param lines of its body, then the
scope its .subckt was defined in, up to the global .param lines); every
{expression} of a line is replaced by its value in the line's scope; inside
instance X1, element R1 becomes R.X1.R1 and internal node n becomes x1.n
(ports replaced by the X line's nodes; ground and .global nodes keep their
names); a .model inside a .subckt is renamed x1.name for that instance.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

MAX_DEPTH = 200       # expression and parameter nesting (netlist/expr.go:19)
MAX_SUBCKT_DEPTH = 64  # nested X instances (netlist/preprocess.go:57)


class ParamError(ValueError):
    """An undefined, circular or non-finite parameter, or a malformed
    expression (netlist.ErrParam)."""


class SubcktError(ValueError):
    """A malformed .subckt or X line (netlist.ErrSubckt)."""


def _go_round_even(v: float) -> float:
    return float(round(v))  # Python's round() rounds half to even, as math.RoundToEven


def _pwr(a):
    return math.copysign(abs(a[0]) ** a[1], a[0])


# Functions an expression may call, with their arity (netlist/expr.go:344-380).
FUNCS = {
    "sqrt": (1, lambda a: math.sqrt(a[0])), "exp": (1, lambda a: math.exp(a[0])),
    "ln": (1, lambda a: math.log(a[0])), "log": (1, lambda a: math.log(a[0])),
    "log10": (1, lambda a: math.log10(a[0])), "sin": (1, lambda a: math.sin(a[0])),
    "cos": (1, lambda a: math.cos(a[0])), "tan": (1, lambda a: math.tan(a[0])),
    "asin": (1, lambda a: math.asin(a[0])), "acos": (1, lambda a: math.acos(a[0])),
    "atan": (1, lambda a: math.atan(a[0])), "atan2": (2, lambda a: math.atan2(a[0], a[1])),
    "sinh": (1, lambda a: math.sinh(a[0])), "cosh": (1, lambda a: math.cosh(a[0])),
    "tanh": (1, lambda a: math.tanh(a[0])), "abs": (1, lambda a: abs(a[0])),
    "floor": (1, lambda a: float(math.floor(a[0]))), "ceil": (1, lambda a: float(math.ceil(a[0]))),
    "int": (1, lambda a: float(math.trunc(a[0]))), "nint": (1, lambda a: _go_round_even(a[0])),
    "sgn": (1, lambda a: 0.0 if a[0] == 0 else math.copysign(1.0, a[0])),
    "min": (2, lambda a: min(a[0], a[1])), "max": (2, lambda a: max(a[0], a[1])),
    "pow": (2, lambda a: a[0] ** a[1]), "pwr": (2, _pwr),
}


def _ident_start(c: str) -> bool:
    return c == "_" or ("a" <= c <= "z") or ("A" <= c <= "Z")


def _ident_char(c: str) -> bool:
    return _ident_start(c) or "0" <= c <= "9"


class _Undefined(Exception):
    pass


class _Expr:
    """Recursive-descent evaluation of one expression (grammar of
    netlist/expr.go:36-47: cond, ||, &&, comparisons, + -, * / %, unary
    - + !, right-associative ^ or **, numbers with SPICE suffixes, names,
    calls and parentheses)."""

    def __init__(self, s: str, scope: Optional["Scope"], depth: int):
        self.s, self.i, self.scope, self.depth, self.nest = s, 0, scope, depth, 0

    def fail(self, msg: str):
        return ParamError(f"expression {self.s!r}: {msg}")

    def space(self):
        while self.i < len(self.s) and self.s[self.i] in " \t":
            self.i += 1

    def eat(self, tok: str) -> bool:
        self.space()
        if self.s.startswith(tok, self.i):
            self.i += len(tok)
            return True
        return False

    def peek(self, tok: str) -> bool:
        self.space()
        return self.s.startswith(tok, self.i)

    def enter(self):
        self.nest += 1
        if self.nest + self.depth > MAX_DEPTH:
            raise self.fail("nested too deeply")

    def cond(self) -> float:
        self.enter()
        try:
            c = self.or_()
            if not self.eat("?"):
                return c
            a = self.cond()
            if not self.eat(":"):
                raise self.fail("missing ':' in c ? a : b")
            b = self.cond()
            return a if c != 0 else b
        finally:
            self.nest -= 1

    def or_(self) -> float:
        v = self.and_()
        while self.eat("||"):
            w = self.and_()
            v = 1.0 if (v != 0 or w != 0) else 0.0
        return v

    def and_(self) -> float:
        v = self.cmp()
        while self.eat("&&"):
            w = self.cmp()
            v = 1.0 if (v != 0 and w != 0) else 0.0
        return v

    def cmp(self) -> float:
        v = self.sum()
        for op in ("==", "!=", "<=", ">=", "<", ">"):
            if self.eat(op):
                w = self.sum()
                r = {"==": v == w, "!=": v != w, "<=": v <= w, ">=": v >= w, "<": v < w, ">": v > w}[op]
                return 1.0 if r else 0.0
        return v

    def sum(self) -> float:
        v = self.prod()
        while True:
            if self.eat("+"):
                v = v + self.prod()
            elif self.eat("-"):
                v = v - self.prod()
            else:
                return v

    def prod(self) -> float:
        v = self.unary()
        while True:
            self.space()
            if self.s.startswith("**", self.i):
                return v  # the power operator belongs to pow
            if self.eat("*"):
                v = v * self.unary()
            elif self.eat("/"):
                w = self.unary()
                if w == 0:
                    raise self.fail("division by zero")
                v = v / w
            elif self.eat("%"):
                w = self.unary()
                if w == 0:
                    raise self.fail("modulo by zero")
                v = math.fmod(v, w)  # math.Mod: the sign of the dividend
            else:
                return v

    def unary(self) -> float:
        self.enter()
        try:
            if self.eat("-"):
                return -self.unary()
            if self.eat("+"):
                return self.unary()
            if self.peek("!") and not self.peek("!="):
                self.eat("!")
                return 1.0 if self.unary() == 0 else 0.0
            return self.pow()
        finally:
            self.nest -= 1

    def pow(self) -> float:
        v = self.atom()
        if self.eat("**") or self.eat("^"):
            w = self.unary()
            try:
                return math.pow(v, w)
            except (OverflowError, ValueError):
                return math.nan  # rejected as non-finite by evaluate
        return v

    def atom(self) -> float:
        from .netlist import parse_value
        self.space()
        if self.i >= len(self.s):
            raise self.fail("unexpected end")
        c = self.s[self.i]
        if c == "(":
            self.i += 1
            v = self.cond()
            if not self.eat(")"):
                raise self.fail("missing ')'")
            return v
        if c.isdigit() or c == ".":
            j = self.i
            while j < len(self.s) and (self.s[j].isdigit() or self.s[j] == "."):
                j += 1
            if j < len(self.s) and self.s[j] in "eE":  # exponent, when digits follow
                k = j + 1
                if k < len(self.s) and self.s[k] in "+-":
                    k += 1
                if k < len(self.s) and self.s[k].isdigit():
                    while k < len(self.s) and self.s[k].isdigit():
                        k += 1
                    j = k
            while j < len(self.s) and _ident_start(self.s[j]):
                j += 1  # engineering suffix and unit letters, as in ParseValue
            try:
                v = parse_value(self.s[self.i:j])
            except ValueError as e:
                raise self.fail(str(e))
            self.i = j
            return v
        if _ident_start(c):
            j = self.i
            while j < len(self.s) and _ident_char(self.s[j]):
                j += 1
            name = self.s[self.i:j].lower()
            self.i = j
            if self.eat("("):
                return self.call(name)
            if self.scope is not None:
                try:
                    return self.scope.value(name, self.depth + self.nest + 1)
                except _Undefined:
                    pass
            if name == "pi":
                return math.pi
            raise ParamError(f"undefined parameter {name!r}")
        raise self.fail(f"unexpected {self.s[self.i:]!r}")

    def call(self, name: str) -> float:
        if name not in FUNCS:
            raise self.fail(f"unknown function {name}()")
        n, f = FUNCS[name]
        args: List[float] = []
        if not self.eat(")"):
            while True:
                args.append(self.cond())
                if self.eat(")"):
                    break
                if not self.eat(","):
                    raise self.fail(f"missing ',' or ')' in {name}()")
        if len(args) != n:
            raise self.fail(f"{name}() takes {n} argument(s), got {len(args)}")
        try:
            return float(f(args))
        except (ValueError, OverflowError, ZeroDivisionError):
            return math.nan  # rejected as non-finite by evaluate


def evaluate(s: str, scope: Optional["Scope"] = None, depth: int = 0) -> float:
    """evalExpr, netlist/expr.go:48-64: the value of s, names resolved in
    scope; the result must be finite."""
    if depth > MAX_DEPTH:
        raise ParamError("expressions nested too deeply")
    p = _Expr(s, scope, depth)
    v = p.cond()
    p.space()
    if p.i < len(p.s):
        raise p.fail(f"unexpected {p.s[p.i:]!r}")
    if math.isnan(v) or math.isinf(v):
        raise ParamError(f"{s!r} is not finite")
    return v


def split_words(s: str) -> Tuple[List[str], List[Tuple[str, str]]]:
    """splitWords, netlist/preprocess.go:156-222: plain words and name=value
    assignments ({braced} or 'quoted' values without their delimiters);
    commas are blanks."""
    words: List[str] = []
    assigns: List[Tuple[str, str]] = []
    i, n = 0, len(s)
    blank = lambda ch: ch in " \t,"
    while True:
        while i < n and blank(s[i]):
            i += 1
        if i >= n:
            return words, assigns
        j = i
        while j < n and not blank(s[j]) and s[j] != "=":
            j += 1
        w = s[i:j]
        i = j
        while i < n and blank(s[i]):
            i += 1
        if i >= n or s[i] != "=":
            if w == "":
                raise ParamError("unexpected '='")
            words.append(w)
            continue
        i += 1
        while i < n and blank(s[i]):
            i += 1
        if not w or not _valid_name(w):
            raise ParamError(f"{w!r} is not a parameter name")
        if i >= n:
            raise ParamError(f"{w}= has no value")
        if s[i] == "{":
            k = _match_brace(s, i)
            v, i = s[i + 1:k], k + 1
        elif s[i] == "'":
            k = s.find("'", i + 1)
            if k < 0:
                raise ParamError(f"unterminated quote in {w}=")
            v, i = s[i + 1:k], k + 1
        else:
            j = i
            while j < n and not blank(s[j]):
                j += 1
            v, i = s[i:j], j
        assigns.append((w.lower(), v))


def _match_brace(s: str, i: int) -> int:
    depth = 0
    for k in range(i, len(s)):
        if s[k] == "{":
            depth += 1
        elif s[k] == "}":
            depth -= 1
            if depth == 0:
                return k
    raise ParamError("unmatched '{'")


def _valid_name(w: str) -> bool:
    return bool(w) and _ident_start(w[0]) and all(_ident_char(c) for c in w[1:])


class Scope:
    """One level of parameters, subcircuit definitions and local models
    (netlist/preprocess.go:245-330)."""

    def __init__(self, parent: Optional["Scope"]):
        self.parent = parent
        self.exprs: Dict[str, str] = {}
        self.vals: Dict[str, float] = {}
        self.busy: set = set()
        self.defs: Dict[str, "Subckt"] = {}
        self.models: Dict[str, str] = {}

    def value(self, name: str, depth: int) -> float:
        c: Optional[Scope] = self
        while c is not None:
            if name in c.vals:
                return c.vals[name]
            if name in c.exprs:
                if name in c.busy:
                    raise ParamError(f"parameter {name} is defined in terms of itself")
                c.busy.add(name)
                try:
                    v = evaluate(c.exprs[name], c, depth)
                finally:
                    c.busy.discard(name)
                c.vals[name] = v
                return v
            c = c.parent
        raise _Undefined(name)

    def define(self, name: str, expr: str):
        if name in self.exprs or name in self.vals:
            raise ParamError(f"parameter {name} defined twice")
        if name in FUNCS:
            raise ParamError(f"{name} is a built-in name")
        self.exprs[name] = expr

    def find_def(self, name: str):
        c: Optional[Scope] = self
        while c is not None:
            if name in c.defs:
                return c.defs[name], c
            c = c.parent
        return None, None

    def model(self, name: str) -> str:
        c: Optional[Scope] = self
        while c is not None:
            if name in c.models:
                return c.models[name]
            c = c.parent
        return name


def substitute(text: str, scope: Scope) -> str:
    """substitute, netlist/preprocess.go:332-366: every {expression} and
    'expression' of text replaced by its value (shortest round-trip
    decimal, as strconv 'g' -1; the element parser reads it back exactly)."""
    if "{" not in text and "'" not in text:
        return text
    out, i = [], 0
    while i < len(text):
        c = text[i]
        if c == "{":
            k = _match_brace(text, i)
            expr, i = text[i + 1:k], k + 1
        elif c == "'":
            k = text.find("'", i + 1)
            if k < 0:
                raise ParamError("unterminated quote")
            expr, i = text[i + 1:k], k + 1
        else:
            out.append(c)
            i += 1
            continue
        out.append(repr(evaluate(expr, scope, 0)))
    return "".join(out)


@dataclass
class Subckt:
    name: str
    ports: List[str]
    defaults: List[Tuple[str, str]]
    body: List[str] = field(default_factory=list)
    defs: Dict[str, "Subckt"] = field(default_factory=dict)


def _head(line: str) -> str:
    f = line.split()
    return f[0].lower() if f else ""


def split_defs(lines: List[str]) -> Tuple[List[str], Dict[str, Subckt]]:
    """splitDefs, netlist/preprocess.go:390-445: the lines outside .subckt
    blocks and the definitions (each with its nested ones)."""
    rest: List[str] = []
    defs: Dict[str, Subckt] = {}
    i = 0
    while i < len(lines):
        ln = lines[i]
        h = _head(ln)
        if h == ".ends":
            raise SubcktError(".ends without .subckt")
        if h != ".subckt":
            rest.append(ln)
            i += 1
            continue
        words, assigns = split_words(ln)
        ports = [w.lower() for w in words[1:] if w.lower() != "params:"]
        if not ports:
            raise SubcktError(".subckt needs: name node [node ...] [params: name=value ...]")
        d = Subckt(ports[0], ports[1:], assigns)
        if len(set(d.ports)) != len(d.ports) or any(p in ("0", "gnd") for p in d.ports):
            raise SubcktError(f".subckt {d.name}: a port is ground or repeated")
        depth, j = 1, i + 1
        while j < len(lines):
            hj = _head(lines[j])
            if hj == ".subckt":
                depth += 1
            elif hj == ".ends":
                depth -= 1
            if depth == 0:
                break
            j += 1
        if depth != 0:
            raise SubcktError(f".subckt {d.name} has no .ends")
        f = lines[j].split()
        if len(f) > 1 and f[1].lower() != d.name:
            raise SubcktError(f".ends {f[1]} closes .subckt {d.name}")
        d.body, d.defs = split_defs(lines[i + 1:j])
        if d.name in defs:
            raise SubcktError(f".subckt {d.name} defined twice")
        defs[d.name] = d
        i = j + 1
    return rest, defs


@dataclass
class _Inst:
    path: str
    ports: Dict[str, str]
    depth: int


class Expander:
    """expander, netlist/preprocess.go:447-720: the flat element and control
    lines, in the order the element parser reads them."""

    def __init__(self):
        self.glob = Scope(None)
        self.globals: set = set()
        self.out: List[str] = []
        self.params: Dict[str, float] = {}

    def run(self, lines: List[str]) -> List[str]:
        for k, ln in enumerate(lines):
            if _head(ln) == ".end":
                lines = lines[:k]
                break
        rest, defs = split_defs(lines)
        self.glob.defs = defs
        self.emit(rest, self.glob, None)
        for n in sorted(self.glob.exprs):  # every global parameter must evaluate
            self.params[n] = self.glob.value(n, 0)
        return self.out

    def collect(self, lines: List[str], s: Scope) -> List[str]:
        rest = []
        for ln in lines:
            h = _head(ln)
            if h == ".param":
                words, assigns = split_words(ln)
                if len(words) != 1 or not assigns:
                    raise ParamError(".param needs: name=value [name=value ...]")
                for name, expr in assigns:
                    s.define(name, expr)
            elif h == ".global":
                for nd in ln.split()[1:]:
                    self.globals.add(nd.lower())
            else:
                rest.append(ln)
        return rest

    def emit(self, lines: List[str], s: Scope, inst: Optional[_Inst]):
        lines = self.collect(lines, s)
        if inst is not None:  # an instance's local .model lines first
            models = [ln for ln in lines if _head(ln) == ".model"]
            lines = models + [ln for ln in lines if _head(ln) != ".model"]
        for ln in lines:
            self.line(ln, s, inst)

    def node(self, n: str, inst: Optional[_Inst]) -> str:
        if inst is None or n in ("0", "gnd") or n in self.globals:
            return n
        if n in inst.ports:
            return inst.ports[n]
        return inst.path.lower() + "." + n

    def line(self, ln: str, s: Scope, inst: Optional[_Inst]):
        head = _head(ln)
        if not head:
            return
        text = ln if head == ".title" else substitute(ln, s)
        if head.startswith("."):
            if inst is None:
                self.out.append(text)
                return
            if head != ".model":
                raise SubcktError(f"{text.split()[0]} is not allowed inside .subckt")
            f = text.split()
            local = inst.path.lower() + "." + f[1]
            s.models[f[1].lower()] = local
            self.out.append(" ".join([f[0], local] + f[2:]))
            return
        if head[0] == "x":
            self.instantiate(text, s, inst)
            return
        if inst is not None:
            text = self.rename(text, s, inst)
        self.out.append(text)

    def rename(self, text: str, s: Scope, inst: _Inst) -> str:
        """rename, netlist/preprocess.go:610-632, for the element kinds this
        path reads: R, C, L, V, I (two nodes), D (two nodes, model), M (four
        nodes, model), Q (three or four nodes, model), E and G (two nodes, two
        control nodes), F and H (two nodes, the controlling source, named as
        the instance names it), K (the two coupled inductors, named so)."""
        f = text.split()
        name = f[0]
        k = name[0].upper()
        nnodes = {"R": 2, "C": 2, "L": 2, "V": 2, "I": 2, "D": 2, "M": 4, "E": 4, "G": 4, "F": 2, "H": 2,
                  "K": 0}.get(k)
        if k == "Q":
            nnodes = 4 if len(f) >= 6 and "=" not in f[5] and not _is_number(f[5]) else 3
        if nnodes is None:
            raise SubcktError(f"element {name} is not on this path inside a .subckt")
        f[0] = name[:1] + "." + inst.path + "." + name
        for q in range(1, 1 + nnodes):
            f[q] = self.node(f[q].lower(), inst)
        if k in ("D", "M", "Q"):
            f[1 + nnodes] = s.model(f[1 + nnodes].lower())
        if k in ("F", "H") and len(f) > 3:
            f[3] = (f[3][:1] + "." + inst.path + "." + f[3]).lower()
        if k == "K":  # el.Couples, netlist/preprocess.go:649-651
            for q in (1, 2):
                if q < len(f):
                    f[q] = (f[q][:1] + "." + inst.path + "." + f[q]).lower()
        return " ".join(f)

    def instantiate(self, text: str, caller: Scope, parent: Optional[_Inst]):
        words, assigns = split_words(text)
        w = [v for v in words if v.lower() != "params:"]
        if len(w) < 2:
            raise SubcktError("X needs: Xname node ... subckt [name=value ...]")
        xname, sub, nodes = w[0], w[-1].lower(), w[1:-1]
        d, dscope = caller.find_def(sub)
        if d is None:
            raise SubcktError(f"{xname}: no .subckt {sub}")
        if len(nodes) != len(d.ports):
            raise SubcktError(f"{xname} has {len(nodes)} nodes, .subckt {d.name} has {len(d.ports)} ports")
        depth, path = (1, xname) if parent is None else (parent.depth + 1, parent.path + "." + xname)
        if depth > MAX_SUBCKT_DEPTH:
            raise SubcktError(f"subcircuits nested deeper than {MAX_SUBCKT_DEPTH} (recursive .subckt?)")
        inst = _Inst(path, {p: self.node(nodes[k].lower(), parent) for k, p in enumerate(d.ports)}, depth)
        s = Scope(dscope)
        s.defs = d.defs
        declared = set()
        for name, expr in d.defaults:
            s.define(name, expr)
            declared.add(name)
        for name, expr in assigns:
            if name not in declared:
                raise ParamError(f"{xname}: .subckt {d.name} has no parameter {name}")
            if name in s.vals:
                raise ParamError(f"{xname}: {name} given twice")
            s.vals[name] = evaluate(expr, caller, 0)
        self.emit(d.body, s, inst)


def _is_number(t: str) -> bool:
    from .netlist import parse_value
    try:
        parse_value(t)
        return True
    except ValueError:
        return False


def expand(lines: List[str]) -> Tuple[List[str], Dict[str, float]]:
    """The flat lines of a netlist body (the lines after the title, with
    comments and continuations already resolved) and its global parameters."""
    x = Expander()
    out = x.run(lines)
    return out, x.params
