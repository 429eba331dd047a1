"""The cut of a JacobianTearing, resolved on the host (NewJacobianTearing and
tearTerminals, kron/jacobian.go:80-170): element names to the node-unknown
pairs whose couplings the torn matrices drop from their pieces, and the cuts
the reference's acceptance test tries on a circuit (jacobianCuts,
kron/jacobian_test.go:23-41)."""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

from . import circuit as kc


class TearError(ValueError):
    """kron.ErrUnsupported: a cut the tearing cannot take."""


def tear_terminals(e: kc.Element) -> List[int]:
    """The non-ground terminals (node numbers) of an element whose Jacobian
    coupling may be torn, each once, in terminal order; elements whose
    coupling runs through a branch-current unknown are refused
    (kron/jacobian.go:136-170)."""
    if e.kind in (kc.V, kc.L, kc.E, kc.F, kc.H):
        raise TearError(f"tearing {e.name}: its coupling runs through a branch-current unknown; "
                        "tear a neighbouring element")
    out: List[int] = []
    for n in tuple(e.nodes) + (tuple(e.ctrl_nodes) if e.kind == kc.G else ()):
        if n != 0 and n not in out:
            out.append(n)
    return out


def cut_pairs(c: kc.Circuit, torn: Sequence[str]) -> np.ndarray:
    """NewJacobianTearing (kron/jacobian.go:95-134): the node-unknown pairs
    (node - 1) of every torn element's terminal pairs, [k, 2] int32."""
    seen = set()
    pairs: List[List[int]] = []
    for name in torn:
        key = name.lower()
        if key in seen:
            raise TearError(f"{name} is torn twice")
        seen.add(key)
        e = next((x for x in c.elements if x.name.lower() == key), None)
        if e is None:
            raise TearError(f"no element {name!r}")
        terms = tear_terminals(e)
        if len(terms) < 2:
            raise TearError(f"tearing {e.name}: it connects to ground, so tearing it separates nothing")
        for x in range(len(terms)):
            for y in range(x + 1, len(terms)):
                if terms[x] != terms[y]:
                    pairs.append([terms[x] - 1, terms[y] - 1])
    return np.asarray(pairs, np.int32).reshape(-1, 2)


def jacobian_cuts(c: kc.Circuit) -> List[List[str]]:
    """Every element that can be torn on its own, then all of them together
    (when there are several)."""
    names = []
    for e in c.elements:
        try:
            if len(tear_terminals(e)) >= 2:
                names.append(e.name)
        except TearError:
            pass
    out = [[n] for n in names]
    if len(names) > 1:
        out.append(names)
    return out
