"""The per-system seam: solver.Matrix (kh_sys_*) and solver.CMatrix
(kh_csys_*, solver/matrix.go:23-46), the matrices analysis gets from
Options.Solver / SolverFactory (analysis/op.go:67-78).

CPU (the library's host side, no device needed):
  * AbsProduct (solver/sparse.go:481-490 real, :519-527 complex): y_i is the
    sum over the stored entries, in first-Add order, of |A_ij|·|x_j| (cmplx.Abs
    magnitudes): equal bit for bit to the oracle's, on systems with repeated
    entries and ground (negative) indices; known answers on a 3x3 complex
    system whose magnitudes are exact (3-4-5 and 5-12-13 triangles).

GPU:
  * the complex seam against the oracle's CSparse (the same Gilbert–Peierls
    with cmplx.Abs pivoting and Go's complex division): solutions bit for
    bit, the same ordering, pivot rows and fill, on complex MNA grids (a
    conductance-capacitance grid at one frequency) and random unsymmetric
    systems; the error classes (singular, non-finite).
"""
import ctypes
import os

import numpy as np
import pytest

from kronspice_amd import api
from oracle import oracle


def hostlib():
    """The library without device initialisation: the seam's host side."""
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libkronhip.so not built")
    L = ctypes.CDLL(api.LIB_PATH)
    P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    for name, args in [("kh_sys_create", [i32, P]), ("kh_sys_add_batch", [P, P, P, P, i64]),
                       ("kh_sys_absproduct", [P, P, P]), ("kh_sys_destroy", [P]),
                       ("kh_csys_create", [i32, P]), ("kh_csys_add_batch", [P, P, P, P, i64]),
                       ("kh_csys_absproduct", [P, P, P]), ("kh_csys_destroy", [P])]:
        getattr(L, name).argtypes = args
    L.kh_sys_destroy.restype = None
    L.kh_csys_destroy.restype = None
    return L


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def stamps(n, m, rng, cplx):
    """m stamps with repeats (about a third revisit an earlier entry) and
    ground indices (-1), as device Load calls produce them."""
    ii, jj, vv = [], [], []
    for k in range(m):
        if k > 3 and rng.random() < 0.35:
            t = rng.integers(0, k)
            i, j = ii[t], jj[t]
        else:
            i, j = int(rng.integers(-1, n)), int(rng.integers(-1, n))
        v = rng.normal() * 10 ** rng.uniform(-3, 3)
        if cplx:
            v = complex(v, rng.normal() * 10 ** rng.uniform(-3, 3))
        ii.append(i); jj.append(j); vv.append(v)
    return np.array(ii, np.int32), np.array(jj, np.int32), np.array(vv, np.complex128 if cplx else np.float64)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_absproduct_matches_oracle(cplx, seed):
    L = hostlib()
    rng = np.random.default_rng(seed)
    n = 40
    i, j, v = stamps(n, 400, rng, cplx)
    h = ctypes.c_void_p()
    create, add, absp, destroy = ((L.kh_csys_create, L.kh_csys_add_batch, L.kh_csys_absproduct, L.kh_csys_destroy)
                                  if cplx else (L.kh_sys_create, L.kh_sys_add_batch, L.kh_sys_absproduct,
                                                L.kh_sys_destroy))
    assert create(n, ctypes.byref(h)) == 0
    assert add(h, ptr(i), ptr(j), ptr(v), len(i)) == 0
    o = oracle.Sparse(n, complex_=cplx)
    for a, b, c in zip(i, j, v):
        o.add(int(a), int(b), c)
    x = rng.normal(size=n) * 10 ** rng.uniform(-2, 2, size=n)
    if cplx:
        x = x + 1j * rng.normal(size=n)
    y = np.zeros(n)
    assert absp(h, ptr(np.ascontiguousarray(x)), ptr(y)) == 0
    destroy(h)
    assert np.array_equal(y, o.absproduct(x))


def test_complex_absproduct_known_answer():
    """y_i = sum_j |A_ij|·|x_j| with exact magnitudes: |3+4i| = 5, |5+12i| = 13,
    |-8-6i| = 10; an entry added twice sums before its magnitude is taken."""
    L = hostlib()
    h = ctypes.c_void_p()
    assert L.kh_csys_create(3, ctypes.byref(h)) == 0
    i = np.array([0, 0, 1, 2, 2, 1, -1], np.int32)
    j = np.array([0, 1, 1, 2, 0, 1, 2], np.int32)
    v = np.array([3 + 4j, 5 + 12j, 1 + 1j, -8 - 6j, 2j, 2 - 1j, 99], np.complex128)  # (1,1) twice: 3 + 0i
    assert L.kh_csys_add_batch(h, ptr(i), ptr(j), ptr(v), len(i)) == 0
    x = np.array([3 + 4j, 0.5, 12 - 5j], np.complex128)  # |x| = 5, 0.5, 13
    y = np.zeros(3)
    assert L.kh_csys_absproduct(h, ptr(x), ptr(y)) == 0
    L.kh_csys_destroy(h)
    assert y.tolist() == [5 * 5 + 13 * 0.5, 3 * 0.5, 10 * 13 + 2 * 5]


def complex_grid(w, h, omega, rng):
    """A w x h grid of conductances with a capacitance to ground at every node
    (Y = G + jwC) and two voltage sources: a complex MNA system."""
    nn = w * h
    n = nn + 2
    st = []

    def y(a, b, v):
        st.extend([(a, a, v), (b, b, v), (a, b, -v), (b, a, -v)])

    for yy in range(h):
        for xx in range(w):
            k = yy * w + xx
            if xx + 1 < w:
                y(k, k + 1, complex(1 / (10 + 990 * rng.random()), 0))
            if yy + 1 < h:
                y(k, k + w, complex(1 / (10 + 990 * rng.random()), 0))
            st.append((k, k, complex(1e-9, omega * 1e-9 * (1 + rng.random()))))
    b = np.zeros(n, np.complex128)
    for s, node in enumerate((0, nn - 1)):
        st.extend([(node, nn + s, 1), (nn + s, node, 1)])
        b[nn + s] = complex(1 + rng.random(), rng.random())
    return n, st, b


def solve_both(n, st, b):
    g = api.CSparse(n)
    o = oracle.Sparse(n, complex_=True)
    for i, j, v in st:
        g.add(i, j, v)
        o.add(i, j, v)
    x = g.solve(b)
    e, xo = o.solve(b)
    return g, o, x, e, xo


@pytest.mark.gpu
@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e3, 2 * np.pi * 1e9])
def test_csys_grid_bitexact_with_oracle(omega):
    rng = np.random.default_rng(7)
    n, st, b = complex_grid(12, 9, omega, rng)
    g, o, x, e, xo = solve_both(n, st, b)
    assert e == 0
    assert np.array_equal(x, xo)
    info = o.info()
    q, prow = g.pivots()
    assert np.array_equal(q, info["q"]) and np.array_equal(prow, info["prow"])
    assert g.nnz() == (info["nnz_a"], info["nnz_l"], info["nnz_u"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_csys_random_unsymmetric_bitexact(seed):
    rng = np.random.default_rng(100 + seed)
    n = 30
    st = [(i, i, complex(rng.normal(), rng.normal())) for i in range(n)]
    for _ in range(4 * n):
        st.append((int(rng.integers(0, n)), int(rng.integers(0, n)), complex(rng.normal(), rng.normal())))
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    g, o, x, e, xo = solve_both(n, st, b)
    assert e == 0
    assert np.array_equal(x, xo)
    q, prow = g.pivots()
    assert np.array_equal(prow, o.info()["prow"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["singular", "nonfinite"])
def test_csys_error_classes(case):
    g = api.CSparse(2)
    if case == "singular":
        g.add(0, 0, 1 + 1j); g.add(1, 0, 2 + 2j)
        want = api.SINGULAR
    else:
        g.add(0, 0, complex(float("nan"), 0)); g.add(1, 1, 1)
        want = api.NONFINITE
    with pytest.raises(api.KhError) as ex:
        g.solve(np.ones(2, np.complex128))
    assert ex.value.status == want
