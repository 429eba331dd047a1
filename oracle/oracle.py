"""ctypes front of the FP64 oracle (liboracle.so).

ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker:
nothing in the product path (kronspice_amd/) uses it.

The library is the CPU restatement of kronspice's Go solve path (see
sparse.hpp, devices.hpp and sim.cpp for the file:line citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ERR_NAMES = {0: "ok", 1: "singular", 2: "nonfinite", 3: "badsize", 4: "noconv", 5: "tstep", 6: "maxsteps",
             8: "ledger", 9: "badspec", 10: "unsupported"}


class KhOptions(C.Structure):
    _fields_ = [("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double), ("gmin", C.c_double),
                ("itl1", C.c_int32), ("itl4", C.c_int32), ("no_gmin_stepping", C.c_int32),
                ("no_source_stepping", C.c_int32), ("no_pseudo_transient", C.c_int32),
                ("skip_energy_check", C.c_int32), ("ledger_tol", C.c_double),
                ("polish", C.c_int32), ("circuits_per_wave", C.c_int32),
                ("staging", C.c_int32), ("waves_per_simd", C.c_int32), ("backend", C.c_int32),
                ("precision", C.c_int32), ("circuit_wave", C.c_int32),
                ("no_schedule", C.c_int32)]  # read by the GPU library only


class KhTranSpec(C.Structure):
    _fields_ = [("step", C.c_double), ("stop", C.c_double), ("start", C.c_double), ("max_step", C.c_double),
                ("uic", C.c_int32), ("method", C.c_int32), ("chgtol", C.c_double), ("trtol", C.c_double),
                ("min_step", C.c_double), ("max_steps", C.c_int64)]


class OpInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("iters", C.c_int32), ("solves", C.c_int64), ("nnz_a", C.c_int64),
                ("nnz_l", C.c_int64), ("nnz_u", C.c_int64), ("analyses", C.c_int32), ("err_step", C.c_int32),
                ("ledger_rel", C.c_double), ("inflow", C.c_double), ("outflow", C.c_double),
                ("dissipation", C.c_double), ("storage", C.c_double)]


class TranInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("accepted", C.c_int32), ("rejected", C.c_int32), ("iters", C.c_int32),
                ("count", C.c_int32), ("min_step", C.c_double), ("max_step", C.c_double), ("max_rel", C.c_double),
                ("t_end", C.c_double)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        _lib.oracle_sparse_new.restype = P
        _lib.oracle_sparse_new.argtypes = [C.c_int32, C.c_int32]
        _lib.oracle_sparse_free.argtypes = [P, C.c_int32]
        _lib.oracle_sparse_clear.argtypes = [P, C.c_int32]
        _lib.oracle_sparse_add.argtypes = [P, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32]
        _lib.oracle_sparse_solve.argtypes = [P, P, P, C.c_int32]
        _lib.oracle_sparse_info.argtypes = [P, C.c_int32, P, P, P, P, P, P]
        _lib.oracle_sparse_absproduct.argtypes = [P, P, P]
        _lib.oracle_csparse_absproduct.argtypes = [P, P, P]
        _lib.oracle_min_degree.argtypes = [C.c_int32, P, P, C.c_int64, P]
        _lib.oracle_dense_solve.argtypes = [C.c_int32, P, P, P, C.c_int64, P, P]
        _lib.oracle_size.argtypes = [C.c_int32, P, C.c_int32, P]
        _lib.oracle_op.argtypes = [C.c_int32, P, C.c_int32, P, P, C.c_int32, P, P, P, P, P, P]
        _lib.oracle_op_polished.argtypes = [C.c_int32, P, C.c_int32, P, P, P, P, P]
        _lib.oracle_op_report.argtypes = [C.c_int32, P, C.c_int32, P, P, C.c_int32, P, P, P]
        _lib.oracle_dc_method.argtypes = [C.c_int32, P, C.c_int32, P, P, C.c_int32, P, P, P, C.c_int32, P]
        _lib.oracle_plant_stepping.argtypes = [C.c_int32, C.c_double]
        _lib.oracle_lte_step.restype = C.c_double
        _lib.oracle_lte_step.argtypes = [C.c_int32, C.c_double, C.c_double, P, P, C.c_int32] + [C.c_double] * 7
        _lib.oracle_tran.argtypes = [C.c_int32, P, C.c_int32, P, P, P, C.c_int32, P, P, P, P, C.c_int32, P]
        _lib.oracle_ac.argtypes = [C.c_int32, P, C.c_int32, P, P, P, C.c_int32, P, P]
        _lib.oracle_ac_ex.argtypes = [C.c_int32, P, C.c_int32, P, P, P, C.c_int32, P, P, P, P]
        _lib.oracle_dc.argtypes = [C.c_int32, P, C.c_int32, P, P, C.c_int32, P, P, P, P, P, P, P, P, P]
        _lib.oracle_batch_op.argtypes = [C.c_int32, P, C.c_int32, C.c_int32, C.c_int32, P, P, C.c_int32,
                                         C.c_int32, P, P, P]
        _lib.oracle_batch_tran.argtypes = [C.c_int32, P, C.c_int32, C.c_int32, C.c_int32, P, P, P, P,
                                           C.c_int32, P, P, P, P, P]
        _lib.oracle_tear_new.restype = P
        _lib.oracle_tear_new.argtypes = [P, C.c_int32, C.c_double]
        _lib.oracle_tear_free.argtypes = [P]
        _lib.oracle_tear_use.argtypes = [P]
        _lib.oracle_tear_stats.argtypes = [P, P]
        _lib.oracle_tmat_new.restype = P
        _lib.oracle_tmat_new.argtypes = [P, C.c_int32, C.c_int32]
        _lib.oracle_tmat_free.argtypes = [P, C.c_int32]
        _lib.oracle_tmat_clear.argtypes = [P, C.c_int32]
        _lib.oracle_tmat_add_batch.argtypes = [P, C.c_int32, P, P, P, C.c_int64]
        _lib.oracle_tmat_solve.argtypes = [P, C.c_int32, P, P]
        _lib.oracle_tmat_absproduct.argtypes = [P, C.c_int32, P, P]
        _lib.oracle_tmat_pieces.argtypes = [P, C.c_int32, P]
        _lib.oracle_tear_record.argtypes = [C.c_int32]
        _lib.oracle_tear_rec_count.restype = C.c_int64
        _lib.oracle_tear_rec_info.argtypes = [C.c_int64, P, P, P, P, P]
        _lib.oracle_tear_rec_get.argtypes = [C.c_int64, P, P, P, P, P]
    return _lib


def _p(a: Optional[np.ndarray]):
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("the oracle takes C-contiguous arrays (pass np.ascontiguousarray of a view)")
    return a.ctypes.data_as(C.c_void_p)


def options(**kw) -> KhOptions:
    o = KhOptions()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Sparse:
    """solver.Sparse / CSparse restatement (solver/sparse.go)."""

    def __init__(self, n: int, complex_: bool = False):
        self.n = n
        self.cplx = 1 if complex_ else 0
        self.h = lib().oracle_sparse_new(n, self.cplx)
        if not self.h:
            raise ValueError("bad size")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_sparse_free(self.h, self.cplx)
            self.h = None

    def clear(self):
        lib().oracle_sparse_clear(self.h, self.cplx)

    def add(self, i: int, j: int, v):
        v = complex(v)
        lib().oracle_sparse_add(self.h, i, j, v.real, v.imag, self.cplx)

    def solve(self, b) -> Tuple[int, np.ndarray]:
        dt = np.complex128 if self.cplx else np.float64
        b = np.ascontiguousarray(b, dtype=dt)
        x = np.zeros(self.n, dtype=dt)
        e = lib().oracle_sparse_solve(self.h, _p(b), _p(x), self.cplx)
        return e, x

    def info(self):
        a, l, u = C.c_int64(), C.c_int64(), C.c_int64()
        q = np.zeros(self.n, np.int32)
        prow = np.zeros(self.n, np.int32)
        es = C.c_int32()
        lib().oracle_sparse_info(self.h, self.cplx, C.byref(a), C.byref(l), C.byref(u), _p(q), _p(prow), C.byref(es))
        return dict(nnz_a=a.value, nnz_l=l.value, nnz_u=u.value, q=q, prow=prow, err_step=es.value)

    def absproduct(self, x) -> np.ndarray:
        y = np.zeros(self.n)
        if self.cplx:
            x = np.ascontiguousarray(x, dtype=np.complex128)
            lib().oracle_csparse_absproduct(self.h, _p(x), _p(y))
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            lib().oracle_sparse_absproduct(self.h, _p(x), _p(y))
        return y


TEAR_STATS = ("solves", "factors", "reuses", "stale", "fallbacks", "refines", "max_pieces")


class Tearing:
    """kron.JacobianTearing restatement (kron/jacobian.go:66-712): the cut
    node pairs (unknown indices, -1 ground) and the counters its matrices add
    to. `with tearing:` makes every matrix the oracle's analyses create inside
    the block a torn one of this tearing (Options.Solver)."""

    def __init__(self, cuts, reuse_tol: float = 0.0):
        c = np.ascontiguousarray(np.asarray(cuts, np.int32).reshape(-1, 2))
        self.cuts = c
        self.h = lib().oracle_tear_new(_p(c) if len(c) else None, len(c), reuse_tol)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_tear_free(self.h)
            self.h = None

    def __enter__(self):
        lib().oracle_tear_use(self.h)
        return self

    def __exit__(self, *exc):
        lib().oracle_tear_use(None)

    def stats(self) -> dict:
        out = np.zeros(7, np.int64)
        lib().oracle_tear_stats(self.h, _p(out))
        return dict(zip(TEAR_STATS, (int(v) for v in out)))

    def matrix(self, n: int, complex_: bool = False) -> "TornMatrix":
        return TornMatrix(self, n, complex_)


class TornMatrix:
    """One torn system (tornMat, kron/jacobian.go:199-589)."""

    def __init__(self, tear: Tearing, n: int, complex_: bool = False):
        self.tear = tear  # keeps the tearing alive
        self.n = n
        self.cplx = 1 if complex_ else 0
        self.h = lib().oracle_tmat_new(tear.h, n, self.cplx)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_tmat_free(self.h, self.cplx)
            self.h = None

    def clear(self):
        lib().oracle_tmat_clear(self.h, self.cplx)

    def add_batch(self, i, j, v):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        v = np.ascontiguousarray(v, np.complex128 if self.cplx else np.float64)
        lib().oracle_tmat_add_batch(self.h, self.cplx, _p(i), _p(j), _p(v), len(i))

    def solve(self, b) -> Tuple[int, np.ndarray]:
        dt = np.complex128 if self.cplx else np.float64
        b = np.ascontiguousarray(b, dtype=dt)
        x = np.zeros(self.n, dtype=dt)
        e = lib().oracle_tmat_solve(self.h, self.cplx, _p(b), _p(x))
        return e, x

    def absproduct(self, x) -> np.ndarray:
        y = np.zeros(self.n)
        x = np.ascontiguousarray(x, np.complex128 if self.cplx else np.float64)
        lib().oracle_tmat_absproduct(self.h, self.cplx, _p(x), _p(y))
        return y

    def pieces(self):
        po = np.zeros(self.n, np.int32)
        k = lib().oracle_tmat_pieces(self.h, self.cplx, _p(po))
        return k, po


class TearRecording:
    """`with TearRecording() as rec:` records every torn solve the oracle
    makes inside the block; rec.records then lists, per solve in call order,
    dict(mat, cplx, n, status, ri, ci, val, b, x) (complex arrays when cplx)."""

    def __enter__(self):
        lib().oracle_tear_record(1)
        self.records = []
        return self

    def __exit__(self, *exc):
        L = lib()
        for k in range(L.oracle_tear_rec_count()):
            mat, cp, n, st = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            ns = C.c_int64()
            L.oracle_tear_rec_info(k, C.byref(mat), C.byref(cp), C.byref(n), C.byref(ns), C.byref(st))
            dt = np.complex128 if cp.value else np.float64
            ri = np.zeros(ns.value, np.int32)
            ci = np.zeros(ns.value, np.int32)
            val = np.zeros(ns.value, dt)
            b = np.zeros(n.value, dt)
            x = np.zeros(n.value, dt)
            L.oracle_tear_rec_get(k, _p(ri), _p(ci), _p(val), _p(b), _p(x))
            self.records.append(dict(mat=mat.value, cplx=cp.value, n=n.value, status=st.value, ri=ri, ci=ci,
                                     val=val, b=b, x=x))
        L.oracle_tear_record(0)


def min_degree(n: int, ri, ci) -> np.ndarray:
    ri = np.ascontiguousarray(ri, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    out = np.zeros(n, np.int32)
    lib().oracle_min_degree(n, _p(ri), _p(ci), len(ri), _p(out))
    return out


def dense_solve(n: int, ri, ci, v, b):
    ri = np.ascontiguousarray(ri, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    e = lib().oracle_dense_solve(n, _p(ri), _p(ci), _p(v), len(v), _p(b), _p(x))
    return e, x


def size(num_nodes: int, els: np.ndarray, params: np.ndarray) -> int:
    return lib().oracle_size(num_nodes, _p(els), len(els), _p(params))


def op(num_nodes: int, els: np.ndarray, params: np.ndarray, opts: Optional[KhOptions] = None,
       backend: Optional[int] = None, guess: Optional[np.ndarray] = None):
    """analysis.OP restatement. Returns (status, x, info dict with q/prow/currents).
    backend (0 auto, 1 dense, 2 sparse) defaults to opts.backend, whose zero
    value is Auto as in analysis.Options."""
    n = size(num_nodes, els, params)
    if n <= 0:
        return -n, None, {}
    x = np.zeros(n)
    q = np.zeros(n, np.int32)
    prow = np.zeros(n, np.int32)
    cur = np.zeros(len(els))
    info = OpInfo()
    o = opts if opts is not None else KhOptions()
    g = None if guess is None else np.ascontiguousarray(guess, np.float64)
    b = o.backend if backend is None else backend
    e = lib().oracle_op(num_nodes, _p(els), len(els), _p(params), C.byref(o), b, _p(g), _p(x), _p(q),
                        _p(prow), _p(cur), C.byref(info))
    d = {f[0]: getattr(info, f[0]) for f in OpInfo._fields_}
    d.update(q=q, prow=prow, current=cur)
    return e, x, d


DC_METHODS = {"newton": 0, "gmin": 1, "source": 2, "ptran": 3}


def dc_method(num_nodes, els, params, method: str, opts=None, trace: bool = False):
    """One DC method alone on a fresh circuit (analysis/newton_test.go:74-104):
    "newton" = dcSolve, "gmin" / "source" / "ptran" = gminStepping,
    sourceStepping, pseudoTransient (analysis/newton.go:95-311).
    Returns (status, x, Newton iterations[, trace]); the trace is an array of
    (method, parameter, status) rows, one per Newton call: method 0 plain,
    1 gmin (the diagonal conductance), 2 source (the source factor),
    3 pseudo-transient (the step h)."""
    n = size(num_nodes, els, params)
    x = np.zeros(n)
    it = C.c_int32()
    o = opts if opts is not None else KhOptions()
    cap = 3 * 20000
    tr = np.zeros(cap) if trace else None
    tl = C.c_int32()
    e = lib().oracle_dc_method(num_nodes, _p(els), len(els), _p(params), C.byref(o), DC_METHODS[method], _p(x),
                               C.byref(it), _p(tr), cap, C.byref(tl))
    if not trace:
        return e, x, it.value
    assert tl.value <= cap
    return e, x, it.value, tr[: tl.value].reshape(-1, 3)


def plant_stepping(which: str, value: float):
    """Planted mistake (tests only): alters one stepping constant of the
    oracle — "gmin_cap" (10), "source_growth" (1.5), "pt_growth" (2); value 0
    restores the reference's."""
    lib().oracle_plant_stepping({"gmin_cap": 1, "source_growth": 2, "pt_growth": 3}[which], value)


def lte_step(method: int, time: float, h: float, ts, qs, hist_n: int, qn: float, i_n: float, i_prev: float,
             abstol: float = 1e-12, reltol: float = 1e-3, chgtol: float = 1e-14, trtol: float = 7.0) -> float:
    """lteStep (device/lte.go:50-85) on an explicit history (most recent first)."""
    ts = np.ascontiguousarray(ts, np.float64)
    qs = np.ascontiguousarray(qs, np.float64)
    return lib().oracle_lte_step(method, time, h, _p(ts), _p(qs), hist_n, qn, i_n, i_prev, abstol, reltol,
                                 chgtol, trtol)


def op_polished(num_nodes, els, params, opts=None, guess=None):
    n = size(num_nodes, els, params)
    x = np.zeros(n)
    it = C.c_int32()
    o = opts if opts is not None else KhOptions()
    g = None if guess is None else np.ascontiguousarray(guess, np.float64)
    e = lib().oracle_op_polished(num_nodes, _p(els), len(els), _p(params), C.byref(o), _p(g), _p(x), C.byref(it))
    return e, x, it.value


REPORT_FIELDS = ("voltage", "current", "power", "storage", "energy", "supplied", "exterior", "channel")
LEDGER_FIELDS = ("inflow", "outflow", "dissipation", "storage_rate", "model_supplied", "hypothetical", "residual",
                 "scale", "roundoff", "relative_residual")


def op_report(num_nodes, els, params, opts=None, polish=False):
    """Sim.result after analysis.OP (analysis/op.go:361-404): returns (status,
    x, rep [nel, 8] in REPORT_FIELDS order, led [10] in LEDGER_FIELDS order);
    rep and led are filled when the operating point exists (status 0 or the
    ledger error)."""
    n = size(num_nodes, els, params)
    x = np.zeros(n)
    rep = np.zeros((len(els), len(REPORT_FIELDS)))
    led = np.zeros(len(LEDGER_FIELDS))
    o = opts if opts is not None else KhOptions()
    e = lib().oracle_op_report(num_nodes, _p(els), len(els), _p(params), C.byref(o), 1 if polish else 0, _p(x),
                               _p(rep), _p(led))
    return e, x, rep, led


def tran(num_nodes: int, els: np.ndarray, params: np.ndarray, spec: KhTranSpec, opts: Optional[KhOptions] = None,
         backend: Optional[int] = None, cap: int = 0):
    n = size(num_nodes, els, params)
    x = np.zeros(n)
    times = np.zeros(max(cap, 1))
    xs = np.zeros((max(cap, 1), n))
    cur = np.zeros((max(cap, 1), len(els)))
    info = TranInfo()
    o = opts if opts is not None else KhOptions()
    b = o.backend if backend is None else backend
    e = lib().oracle_tran(num_nodes, _p(els), len(els), _p(params), C.byref(o), C.byref(spec), b, _p(x),
                          _p(times), _p(xs), _p(cur), cap, C.byref(info))
    d = {f[0]: getattr(info, f[0]) for f in TranInfo._fields_}
    c = info.count
    d.update(times=times[:c], xs=xs[:c], current=cur[:c])
    return e, x, d


def sweep_values(start: float, stop: float, step: float) -> np.ndarray:
    """DCSweep.values (analysis/dc.go:31-46): start to stop inclusive by step."""
    cnt = (stop - start) / step
    count = int(np.floor(cnt + 1e-9)) + 1
    return start + np.arange(count, dtype=np.float64) * step


def dc(num_nodes: int, els: np.ndarray, params: np.ndarray, sweeps, opts: Optional[KhOptions] = None,
       currents: bool = False):
    """analysis.DC restatement. sweeps: [(element index, start, stop, step)], inner
    first (one or two). Returns (status, x [points, n], iters [points], points done,
    max ledger residual[, element currents [points, nel]])."""
    n = size(num_nodes, els, params)
    total = 1
    for _, a, b, s in sweeps:
        total *= len(sweep_values(a, b, s))
    x = np.zeros((total, n))
    it = np.zeros(total, np.int32)
    src = np.array([s[0] for s in sweeps], np.int32)
    st = np.array([s[1] for s in sweeps], np.float64)
    sp = np.array([s[2] for s in sweeps], np.float64)
    sd = np.array([s[3] for s in sweeps], np.float64)
    done = C.c_int32()
    mr = C.c_double()
    o = opts if opts is not None else KhOptions()
    cur = np.zeros((total, len(els))) if currents else None
    e = lib().oracle_dc(num_nodes, _p(els), len(els), _p(np.ascontiguousarray(params, np.float64)), C.byref(o),
                        len(sweeps), _p(src), _p(st), _p(sp), _p(sd), _p(x), _p(it), C.byref(done), C.byref(mr),
                        _p(cur))
    if currents:
        return e, x, it, done.value, mr.value, cur
    return e, x, it, done.value, mr.value


def ac(num_nodes: int, els: np.ndarray, params: np.ndarray, freqs, opts: Optional[KhOptions] = None,
       currents: bool = False):
    """analysis.AC restatement: (status, X [nf, n] complex, max ledger residual[,
    element currents [nf, nel] complex, ledgers [nf, 6]: inflow, outflow,
    dissipation, reactive, model-supplied, residual])."""
    n = size(num_nodes, els, params)
    freqs = np.ascontiguousarray(freqs, np.float64)
    out = np.zeros((len(freqs), n), np.complex128)
    mr = C.c_double()
    o = opts if opts is not None else KhOptions()
    if not currents:
        e = lib().oracle_ac(num_nodes, _p(els), len(els), _p(params), C.byref(o), _p(freqs), len(freqs), _p(out),
                            C.byref(mr))
        return e, out, mr.value
    cur = np.zeros((len(freqs), len(els)), np.complex128)
    led = np.zeros((len(freqs), 6))
    e = lib().oracle_ac_ex(num_nodes, _p(els), len(els), _p(params), C.byref(o), _p(freqs), len(freqs), _p(out),
                           C.byref(mr), _p(cur), _p(led))
    return e, out, mr.value, cur, led


def batch_op(num_nodes: int, els: np.ndarray, params_soa: np.ndarray, n: int, opts=None, threads: int = 1,
             polish: bool = False):
    """OP over B circuits of one topology; params [P, B]; returns x [n, B], status, iters."""
    P, B = params_soa.shape
    params_soa = np.ascontiguousarray(params_soa, np.float64)
    x = np.zeros((n, B))
    st = np.zeros(B, np.int32)
    it = np.zeros(B, np.int32)
    o = opts if opts is not None else KhOptions()
    lib().oracle_batch_op(num_nodes, _p(els), len(els), P, B, _p(params_soa), C.byref(o), threads,
                          1 if polish else 0, _p(x), _p(st), _p(it))
    return x, st, it


def batch_tran(num_nodes: int, els: np.ndarray, params_soa: np.ndarray, n: int, spec: KhTranSpec, opts=None,
               threads: int = 1, lane_spec: Optional[np.ndarray] = None):
    P, B = params_soa.shape
    params_soa = np.ascontiguousarray(params_soa, np.float64)
    x = np.zeros((n, B))
    st = np.zeros(B, np.int32)
    acc = np.zeros(B, np.int32)
    rej = np.zeros(B, np.int32)
    led = np.zeros(B)
    o = opts if opts is not None else KhOptions()
    ls = None if lane_spec is None else np.ascontiguousarray(lane_spec, np.float64)
    lib().oracle_batch_tran(num_nodes, _p(els), len(els), P, B, _p(params_soa), C.byref(o), C.byref(spec), _p(ls),
                            threads, _p(x), _p(st), _p(acc), _p(rej), _p(led))
    return x, st, acc, rej, led
