"""Python front of libkronhip.so (include/kronhip.h).

The host-side mirror of kronspice's solve-path interfaces:
  * `Sparse` — solver.Matrix (solver/matrix.go:23-36) backed by the GPU:
    Clear / Add / Solve / AbsProduct, pivots and fill as solver.Sparse.
  * `LUBatch` — numeric Gilbert–Peierls LU of B systems sharing one pattern.
  * `op` / `tran` — analysis.OP / analysis.Tran (analysis/op.go:331,
    analysis/tran.go:70) for a batch of circuits of one topology, with the
    reference's Options / TranOptions fields and its sentinel errors as
    per-circuit status codes.

Device memory and streams come from PyTorch (ROCm); every computation runs
in the HIP library. There is no CPU fallback: when the library or the GPU is
missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkronhip.so")

# Status codes (include/kronhip.h), named as the reference's sentinel errors.
OK, SINGULAR, NONFINITE, BADSIZE, NOCONV, TSTEP_TOO_SMALL, TOO_MANY_STEPS = 0, 1, 2, 3, 4, 5, 6
INTERRUPTED, LEDGER, BAD_TRAN_SPEC, UNSUPPORTED, BAD_ARGUMENT, HIP_ERROR, NO_DEVICE, CAPACITY = 7, 8, 9, 10, 11, 12, 13, 14


class KhError(RuntimeError):
    def __init__(self, status: int, detail: str):
        super().__init__(f"kronhip status {status}: {detail}")
        self.status = status


class Options(C.Structure):
    """analysis.Options (analysis/op.go:22-62); zero values take the SPICE defaults."""
    _fields_ = [("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double), ("gmin", C.c_double),
                ("itl1", C.c_int32), ("itl4", C.c_int32), ("no_gmin_stepping", C.c_int32),
                ("no_source_stepping", C.c_int32), ("no_pseudo_transient", C.c_int32),
                ("skip_energy_check", C.c_int32), ("ledger_tol", C.c_double),
                ("polish", C.c_int32), ("circuits_per_wave", C.c_int32),
                ("staging", C.c_int32), ("waves_per_simd", C.c_int32),
                # Options.Backend: 0 auto (dense below 30 unknowns), 1 dense, 2 sparse
                ("backend", C.c_int32),
                # sparse LU arithmetic: 0 FP64, 1 FP32 factors + FP64 refinement
                ("precision", C.c_int32), ("circuit_wave", C.c_int32),
                ("no_schedule", C.c_int32)]


class TranSpec(C.Structure):
    """TranSpec + TranOptions (analysis/tran.go:26-59)."""
    _fields_ = [("step", C.c_double), ("stop", C.c_double), ("start", C.c_double), ("max_step", C.c_double),
                ("uic", C.c_int32), ("method", C.c_int32), ("chgtol", C.c_double), ("trtol", C.c_double),
                ("min_step", C.c_double), ("max_steps", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("solves", C.c_int64), ("newton_iterations", C.c_int64), ("accepted_steps", C.c_int64),
                ("rejected_steps", C.c_int64), ("analyses", C.c_int64), ("patterns", C.c_int64),
                ("ticks", C.c_int64), ("bytes_algorithmic", C.c_int64), ("seconds", C.c_double),
                ("lu_seconds", C.c_double), ("control_seconds", C.c_double), ("lu_launches", C.c_int64),
                ("control_launches", C.c_int64), ("cycles_lu", C.c_double), ("cycles_control", C.c_double),
                ("refine_passes", C.c_int64), ("fp64_redos", C.c_int64),
                ("cycles_stamp", C.c_double), ("cycles_converge", C.c_double), ("cycles_lte", C.c_double),
                ("cycles_accept", C.c_double), ("cycles_ledger", C.c_double),
                ("staged_bytes", C.c_int64), ("staged_partial", C.c_int64),
                ("sched_solves", C.c_int64), ("sched_redos", C.c_int64), ("sched_probes", C.c_int64),
                ("sched_rounds", C.c_int64), ("sched_host_seconds", C.c_double), ("sched_seconds", C.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


_lib = None


def lib():
    """Load libkronhip.so; raises if it was not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # PyTorch ships its own HIP runtime with the same soname
        # (libamdhip64.so.7). Loading torch first makes the library bind to that
        # runtime, so both share one device context and one set of streams.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        P, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.kh_init.restype = i32
        L.kh_last_error.restype = C.c_char_p
        L.kh_status_string.restype = C.c_char_p
        L.kh_status_string.argtypes = [i32]
        L.kh_sys_create.argtypes = [i32, P]
        L.kh_sys_clear.argtypes = [P]
        L.kh_sys_add_batch.argtypes = [P, P, P, P, i64]
        L.kh_sys_solve.argtypes = [P, P, P]
        L.kh_sys_solve_static.argtypes = [P, P, P]
        L.kh_sys_absproduct.argtypes = [P, P, P]
        L.kh_sys_nnz.argtypes = [P, P, P, P]
        L.kh_sys_pivots.argtypes = [P, P, P]
        L.kh_sys_destroy.argtypes = [P]
        L.kh_sys_destroy.restype = None
        for name in ("kh_csys_clear", "kh_csys_destroy"):
            getattr(L, name).argtypes = [P]
        L.kh_csys_destroy.restype = None
        L.kh_csys_create.argtypes = [i32, P]
        L.kh_csys_add_batch.argtypes = [P, P, P, P, i64]
        L.kh_csys_solve.argtypes = [P, P, P]
        L.kh_csys_absproduct.argtypes = [P, P, P]
        L.kh_csys_nnz.argtypes = [P, P, P, P]
        L.kh_csys_pivots.argtypes = [P, P, P]
        L.kh_lu_batch_create.argtypes = [i32, i64, P, P, i32, P]
        L.kh_lu_batch_solve.argtypes = [P, P, P, P, P, P, P]
        L.kh_lu_batch_solve_mixed.argtypes = [P, P, P, P, P, P, P, P]
        L.kh_lu_batch_info.argtypes = [P, P, P, P, P, P]
        L.kh_lu_batch_destroy.argtypes = [P]
        L.kh_tear_create.argtypes = [P, i32, C.c_double, P]
        L.kh_tear_get_stats.argtypes = [P, P]
        L.kh_tear_destroy.argtypes = [P]
        L.kh_tear_destroy.restype = None
        L.kh_tmat_create.argtypes = [P, i32, i32, P]
        L.kh_tmat_clear.argtypes = [P]
        L.kh_tmat_add_batch.argtypes = [P, P, P, P, i64]
        L.kh_tmat_solve.argtypes = [P, P, P]
        L.kh_tmat_absproduct.argtypes = [P, P, P]
        L.kh_tmat_pieces.argtypes = [P, P, P]
        L.kh_tmat_destroy.argtypes = [P]
        L.kh_tmat_destroy.restype = None
        L.kh_tear_batch_create.argtypes = [P, i32, i64, P, P, i32, i32, P]
        L.kh_tear_batch_solve.argtypes = [P, P, P, P, P, P]
        L.kh_tear_batch_pieces.argtypes = [P, P, P]
        L.kh_tear_batch_destroy.argtypes = [P]
        L.kh_tear_batch_destroy.restype = None
        L.kh_lu_batch_destroy.restype = None
        L.kh_topology_create.argtypes = [i32, P, i32, i32, P, P]
        L.kh_topology_size.argtypes = [P]
        L.kh_topology_destroy.argtypes = [P]
        L.kh_topology_destroy.restype = None
        L.kh_batch_run.argtypes = [i32, P, P, P, P]
        L.kh_candidate_size.argtypes = [P]
        L.kh_candidate_size.restype = i32
        L.kh_generation_run.argtypes = [i32, P, P, P, P, P, P, P, P, P]
        L.kh_batch_op.argtypes = [P, i32, P, P, P, P, P, P]
        L.kh_batch_tran.argtypes = [P, i32, P, P, P, P, P, P, P, P, P, P]
        L.kh_batch_dc.argtypes = [P, i32, P, P, P, P, P, P, P, P, P, P]
        L.kh_batch_ac.argtypes = [P, i32, P, P, P, i32, P, P, P, P, P, P]
        _lib = L
        st = L.kh_init()
        if st != OK:
            raise KhError(st, L.kh_last_error().decode())
    return _lib


def _check(st: int):
    if st != OK:
        raise KhError(st, lib().kh_last_error().decode())


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())  # torch tensor on the device


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible to PyTorch: the solve path runs only on the GPU")
    return torch


def status_name(s: int) -> str:
    return lib().kh_status_string(int(s)).decode()


# ------------------------------------------------------------------------------ solver seam
class Sparse:
    """solver.Matrix on the GPU: stamps are staged on the host and cross into
    the library once per Solve (kh_sys_add_batch)."""

    def __init__(self, n: int):
        self.n = n
        self.h = C.c_void_p()
        _check(lib().kh_sys_create(n, C.byref(self.h)))
        self._i, self._j, self._v = [], [], []

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_sys_destroy(self.h)
            self.h = None

    def clear(self):
        self._flush()
        _check(lib().kh_sys_clear(self.h))

    def add(self, i: int, j: int, v: float):
        self._i.append(i); self._j.append(j); self._v.append(v)

    def _flush(self):
        if self._i:
            i = np.asarray(self._i, np.int32); j = np.asarray(self._j, np.int32); v = np.asarray(self._v, np.float64)
            _check(lib().kh_sys_add_batch(self.h, _p(i), _p(j), _p(v), len(i)))
            self._i, self._j, self._v = [], [], []

    def solve(self, b) -> np.ndarray:
        self._flush()
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        _check(lib().kh_sys_solve(self.h, _p(b), _p(x)))
        return x

    def solve_static(self, b) -> np.ndarray:
        """solve without device allocation: the library's static device array
        holds the system (32 MiB per GPU; larger systems raise KhError BADSIZE;
        overlapping calls on one GPU race)."""
        self._flush()
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        _check(lib().kh_sys_solve_static(self.h, _p(b), _p(x)))
        return x

    def abs_product(self, x) -> np.ndarray:
        self._flush()
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        _check(lib().kh_sys_absproduct(self.h, _p(x), _p(y)))
        return y

    def nnz(self):
        a, l, u = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().kh_sys_nnz(self.h, C.byref(a), C.byref(l), C.byref(u)))
        return a.value, l.value, u.value

    def pivots(self):
        q = np.zeros(self.n, np.int32); prow = np.zeros(self.n, np.int32)
        _check(lib().kh_sys_pivots(self.h, _p(q), _p(prow)))
        return q, prow


class CSparse:
    """solver.CMatrix on the GPU (CSparse, solver/sparse.go:495-531): complex
    stamps staged on the host, one kh_csys_add_batch per Solve."""

    def __init__(self, n: int):
        self.n = n
        self.h = C.c_void_p()
        _check(lib().kh_csys_create(n, C.byref(self.h)))
        self._i, self._j, self._v = [], [], []

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_csys_destroy(self.h)
            self.h = None

    def clear(self):
        self._flush()
        _check(lib().kh_csys_clear(self.h))

    def add(self, i: int, j: int, v: complex):
        self._i.append(i); self._j.append(j); self._v.append(complex(v))

    def _flush(self):
        if self._i:
            i = np.asarray(self._i, np.int32); j = np.asarray(self._j, np.int32)
            v = np.asarray(self._v, np.complex128)
            _check(lib().kh_csys_add_batch(self.h, _p(i), _p(j), _p(v), len(i)))
            self._i, self._j, self._v = [], [], []

    def solve(self, b) -> np.ndarray:
        self._flush()
        b = np.ascontiguousarray(b, np.complex128)
        x = np.zeros(self.n, np.complex128)
        _check(lib().kh_csys_solve(self.h, _p(b), _p(x)))
        return x

    def abs_product(self, x) -> np.ndarray:
        self._flush()
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.n)
        _check(lib().kh_csys_absproduct(self.h, _p(x), _p(y)))
        return y

    def nnz(self):
        a, l, u = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().kh_csys_nnz(self.h, C.byref(a), C.byref(l), C.byref(u)))
        return a.value, l.value, u.value

    def pivots(self):
        q = np.zeros(self.n, np.int32); prow = np.zeros(self.n, np.int32)
        _check(lib().kh_csys_pivots(self.h, _p(q), _p(prow)))
        return q, prow


# ------------------------------------------------------------------------------ JacobianTearing
class TearStats(C.Structure):
    _fields_ = [("solves", C.c_int64), ("factors", C.c_int64), ("reuses", C.c_int64), ("stale_reuses", C.c_int64),
                ("fallbacks", C.c_int64), ("refinements", C.c_int64), ("max_pieces", C.c_int64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class JacobianTearing:
    """kron.JacobianTearing on the GPU (kron/jacobian.go:66-712): a
    SolverFactory whose matrices solve by tearing along `cuts` (node-unknown
    pairs, -1 ground, as kronspice_amd.tear.cut_pairs resolves them from
    element names), reusing unchanged piece factors (or those within
    reuse_tol) between solves."""

    def __init__(self, cuts, reuse_tol: float = 0.0):
        c = np.ascontiguousarray(np.asarray(cuts, np.int32).reshape(-1, 2))
        self.h = C.c_void_p()
        _check(lib().kh_tear_create(_p(c) if len(c) else None, len(c), float(reuse_tol), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_tear_destroy(self.h)
            self.h = None

    def stats(self) -> dict:
        s = TearStats()
        _check(lib().kh_tear_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def matrix(self, n: int) -> "TornMatrix":
        return TornMatrix(self, n, False)

    def cmatrix(self, n: int) -> "TornMatrix":
        return TornMatrix(self, n, True)

    def batch(self, n: int, ri, ci, B: int, complex_: bool = False) -> "TearBatch":
        return TearBatch(self, n, ri, ci, B, complex_)


class TornMatrix:
    """JacobianTearing.Matrix / CMatrix: stamps staged on the host, one
    kh_tmat_add_batch per Solve; the torn solve runs on the GPU."""

    def __init__(self, jt: JacobianTearing, n: int, complex_: bool):
        self.jt = jt  # the tearing outlives its matrices
        self.n = n
        self.cplx = complex_
        self.dt = np.complex128 if complex_ else np.float64
        self.h = C.c_void_p()
        _check(lib().kh_tmat_create(jt.h, n, 1 if complex_ else 0, C.byref(self.h)))
        self._i, self._j, self._v = [], [], []

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_tmat_destroy(self.h)
            self.h = None

    def clear(self):
        self._flush()
        _check(lib().kh_tmat_clear(self.h))

    def add(self, i: int, j: int, v):
        self._i.append(i); self._j.append(j); self._v.append(v)

    def add_batch(self, i, j, v):
        self._flush()
        i = np.ascontiguousarray(i, np.int32); j = np.ascontiguousarray(j, np.int32)
        v = np.ascontiguousarray(v, self.dt)
        _check(lib().kh_tmat_add_batch(self.h, _p(i), _p(j), _p(v), len(i)))

    def _flush(self):
        if self._i:
            i, j, v = self._i, self._j, self._v
            self._i, self._j, self._v = [], [], []
            self.add_batch(i, j, v)

    def solve(self, b) -> np.ndarray:
        self._flush()
        b = np.ascontiguousarray(b, self.dt)
        x = np.zeros(self.n, self.dt)
        _check(lib().kh_tmat_solve(self.h, _p(b), _p(x)))
        return x

    def abs_product(self, x) -> np.ndarray:
        self._flush()
        x = np.ascontiguousarray(x, self.dt)
        y = np.zeros(self.n)
        _check(lib().kh_tmat_absproduct(self.h, _p(x), _p(y)))
        return y

    def pieces(self):
        k = C.c_int32()
        po = np.zeros(self.n, np.int32)
        _check(lib().kh_tmat_pieces(self.h, C.byref(k), _p(po)))
        return k.value, po


class TearBatch:
    """B torn systems of one pattern (entries ri, ci in first-Add order), one
    per lane, each keeping its piece factors between solves."""

    def __init__(self, jt: JacobianTearing, n: int, ri, ci, B: int, complex_: bool = False):
        self.jt = jt
        self.n, self.B, self.cplx = n, B, complex_
        ri = np.ascontiguousarray(ri, np.int32); ci = np.ascontiguousarray(ci, np.int32)
        self.nslots = len(ri)
        self.h = C.c_void_p()
        _check(lib().kh_tear_batch_create(jt.h, n, self.nslots, _p(ri), _p(ci), B, 1 if complex_ else 0,
                                          C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_tear_batch_destroy(self.h)
            self.h = None

    def solve(self, vals, b, x, status, stream=None):
        """vals [nslots, B], b and x [n, B] device tensors (float64, or
        complex128 when complex); status [B] int32."""
        _check(lib().kh_tear_batch_solve(self.h, _p(vals), _p(b), _p(x), _p(status),
                                         None if stream is None else C.c_void_p(stream)))

    def pieces(self):
        k = C.c_int32()
        po = np.zeros(self.n, np.int32)
        _check(lib().kh_tear_batch_pieces(self.h, C.byref(k), _p(po)))
        return k.value, po


# ------------------------------------------------------------------------------ batched LU
class LUBatch:
    """B systems with one entry list (ri, ci); values SoA [nnz, B] on the device."""

    def __init__(self, n: int, ri, ci, B: int):
        self.n, self.B = n, B
        ri = np.ascontiguousarray(ri, np.int32); ci = np.ascontiguousarray(ci, np.int32)
        self.nnz = len(ri)
        self.h = C.c_void_p()
        _check(lib().kh_lu_batch_create(n, self.nnz, _p(ri), _p(ci), B, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_lu_batch_destroy(self.h)
            self.h = None

    def solve(self, vals, b, x, status=None, prow=None, stream=None):
        """vals [nnz, B], b and x [n, B] float64 device tensors; status [B] int32."""
        _check(lib().kh_lu_batch_solve(self.h, _p(vals), _p(b), _p(x), _p(status), _p(prow),
                                       C.c_void_p(stream) if stream else None))

    def solve_mixed(self, vals, b, x, status=None, prow=None, passes=None, stream=None):
        """FP32 factors + FP64 refinement; passes [B] int32 receives the FP32
        solves per system (0 where the system was redone in FP64)."""
        _check(lib().kh_lu_batch_solve_mixed(self.h, _p(vals), _p(b), _p(x), _p(status), _p(prow), _p(passes),
                                             C.c_void_p(stream) if stream else None))

    def info(self):
        q = np.zeros(self.n, np.int32)
        l, u, cl, cu = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().kh_lu_batch_info(self.h, _p(q), C.byref(l), C.byref(u), C.byref(cl), C.byref(cu)))
        return dict(q=q, nnz_l=l.value, nnz_u=u.value, cap_l=cl.value, cap_u=cu.value)


# ------------------------------------------------------------------------------ analyses
class Topology:
    """One circuit topology (element table + parameter layout) registered with the library."""

    def __init__(self, num_nodes: int, elements: np.ndarray, params_c0: np.ndarray):
        self.num_nodes = num_nodes
        self.elements = np.ascontiguousarray(elements)
        p0 = np.ascontiguousarray(params_c0, np.float64)
        self.nparams = len(p0)
        self.n_elems = len(self.elements)
        self.h = C.c_void_p()
        _check(lib().kh_topology_create(num_nodes, _p(self.elements), len(self.elements), self.nparams, _p(p0),
                                        C.byref(self.h)))
        self.n = lib().kh_topology_size(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().kh_topology_destroy(self.h)
            self.h = None


@dataclass
class Result:
    x: object          # [n, B] device tensor
    status: object     # [B] int32 device tensor
    iters: object = None
    accepted: object = None
    rejected: object = None
    ledger: object = None
    stats: dict = field(default_factory=dict)
    points: object = None  # DC sweep: [B] points solved
    report: object = None         # [B, n_elems, len(REPORT_FIELDS)] device tensor (op(report=True))
    report_ledger: object = None  # [B, len(LEDGER_FIELDS)] device tensor


def to_device(a: np.ndarray, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to(device="cuda", dtype=dtype) if dtype is not None else t.to("cuda")


def op(topo: Topology, params, opts: Optional[Options] = None, polish: bool = False, nodeset=None,
       report: bool = False) -> Result:
    """analysis.OP for every circuit; params [P, B] (device tensor or numpy).
    nodeset [numNodes-1, B] (optional): Newton's starting node voltages per
    circuit (Options.NodeSet, analysis/op.go:39-43). report: also the
    output.Result entries, Result.report [B, n_elems, len(REPORT_FIELDS)] and
    Result.report_ledger [B, len(LEDGER_FIELDS)]."""
    torch = _torch()
    if isinstance(params, np.ndarray):
        params = to_device(params)
    if isinstance(nodeset, np.ndarray):
        nodeset = to_device(nodeset)
    B = params.shape[1]
    x = torch.zeros((topo.n, B), dtype=torch.float64, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    it = torch.zeros(B, dtype=torch.int32, device="cuda")
    o = Options() if opts is None else opts
    o.polish = 1 if polish else o.polish
    s = Stats()
    torch.cuda.synchronize()
    rep = led = None
    if nodeset is None and not report:
        _check(lib().kh_batch_op(topo.h, B, _p(params), C.byref(o), _p(x), _p(st), _p(it), C.byref(s)))
    else:
        g = Group(topology=topo.h.value, B=B, d_params=params.data_ptr(), d_x=x.data_ptr(), d_status=st.data_ptr(),
                  d_iters=it.data_ptr())
        if nodeset is not None:
            if tuple(nodeset.shape) != (topo.num_nodes - 1, B):
                raise ValueError(f"nodeset shape {tuple(nodeset.shape)}, want {(topo.num_nodes - 1, B)}")
            g.d_nodeset = nodeset.data_ptr()
        if report:
            rep = torch.zeros((B, topo.n_elems, len(REPORT_FIELDS)), dtype=torch.float64, device="cuda")
            led = torch.zeros((B, len(LEDGER_FIELDS)), dtype=torch.float64, device="cuda")
            g.d_report, g.d_report_ledger = rep.data_ptr(), led.data_ptr()
        _check(lib().kh_batch_run(1, C.byref(g), C.byref(o), None, C.byref(s)))
    r = Result(x, st, iters=it, stats=s.as_dict())
    r.report, r.report_ledger = rep, led
    return r


class DcSpec(C.Structure):
    """DCSweep x 1 or 2 (analysis/dc.go:22-28); element = index of the swept source."""
    _fields_ = [("nsweeps", C.c_int32), ("element", C.c_int32 * 2), ("reserved", C.c_int32),
                ("start", C.c_double * 2), ("stop", C.c_double * 2), ("step", C.c_double * 2)]


def dc_spec(sweeps) -> DcSpec:
    """sweeps: [(element index, start, stop, step)], inner first."""
    s = DcSpec()
    s.nsweeps = len(sweeps)
    for k, (e, a, b, h) in enumerate(sweeps):
        s.element[k], s.start[k], s.stop[k], s.step[k] = e, a, b, h
    return s


def sweep_count(start: float, stop: float, step: float) -> int:
    """Number of points of DCSweep.values (analysis/dc.go:31-46)."""
    return int(np.floor((stop - start) / step + 1e-9)) + 1


class DcOutputs(C.Structure):
    _fields_ = [("d_sweep_x", C.c_void_p), ("d_sweep_i", C.c_void_p), ("d_status", C.c_void_p),
                ("d_points", C.c_void_p), ("d_iters", C.c_void_p), ("d_ledger", C.c_void_p)]


def dc(topo: Topology, params, sweeps, opts: Optional[Options] = None, nodeset=None, currents: bool = False) -> Result:
    """analysis.DC for every circuit; sweeps [(element index, start, stop, step)],
    inner first. Result.x is [points, n, B]; iters [points, B] Newton iterations per
    point; points [B] points solved; ledger [B] largest ledger residual; with
    currents, Result.currents [points, n_elems, B] each element's current at
    each point (DCSweepResult.I)."""
    torch = _torch()
    if isinstance(params, np.ndarray):
        params = to_device(params)
    if isinstance(nodeset, np.ndarray):
        nodeset = to_device(nodeset)
    B = params.shape[1]
    pts = 1
    for _, a, b, h in sweeps:
        pts *= sweep_count(a, b, h)
    x = torch.zeros((pts, topo.n, B), dtype=torch.float64, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    done = torch.zeros(B, dtype=torch.int32, device="cuda")
    it = torch.zeros((pts, B), dtype=torch.int32, device="cuda")
    led = torch.zeros(B, dtype=torch.float64, device="cuda")
    o = Options() if opts is None else opts
    spec = dc_spec(sweeps)
    s = Stats()
    torch.cuda.synchronize()
    if not currents:
        _check(lib().kh_batch_dc(topo.h, B, _p(params), C.byref(o), C.byref(spec), _p(nodeset), _p(x), _p(st),
                                 _p(done), _p(it), _p(led), C.byref(s)))
        return Result(x, st, iters=it, ledger=led, stats=s.as_dict(), points=done)
    cur = torch.zeros((pts, topo.n_elems, B), dtype=torch.float64, device="cuda")
    out = DcOutputs(d_sweep_x=x.data_ptr(), d_sweep_i=cur.data_ptr(), d_status=st.data_ptr(),
                    d_points=done.data_ptr(), d_iters=it.data_ptr(), d_ledger=led.data_ptr())
    _check(lib().kh_batch_dc_ex(topo.h, B, _p(params), C.byref(o), C.byref(spec), _p(nodeset), C.byref(out),
                                C.byref(s)))
    r = Result(x, st, iters=it, ledger=led, stats=s.as_dict(), points=done)
    r.currents = cur
    return r


class AcOutputs(C.Structure):
    _fields_ = [("d_xac", C.c_void_p), ("d_iac", C.c_void_p), ("d_ledger_f", C.c_void_p), ("d_status", C.c_void_p),
                ("d_points", C.c_void_p), ("d_ledger", C.c_void_p)]


def ac(topo: Topology, params, freqs, opts: Optional[Options] = None, nodeset=None, currents: bool = False) -> Result:
    """analysis.AC for every circuit at the host frequencies freqs [nf].
    Result.x is a complex128 tensor [nf, n, B]; points [B] frequencies solved;
    ledger [B] largest ledger residual; with currents, Result.currents
    [nf, n_elems, B] complex element currents and Result.ledgers [nf, 6, B]
    (inflow, outflow, dissipation, reactive, model-supplied, residual)."""
    torch = _torch()
    if isinstance(params, np.ndarray):
        params = to_device(params)
    if isinstance(nodeset, np.ndarray):
        nodeset = to_device(nodeset)
    B = params.shape[1]
    f = np.ascontiguousarray(freqs, np.float64)
    nf = len(f)
    x = torch.zeros((nf, topo.n, B, 2), dtype=torch.float64, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    done = torch.zeros(B, dtype=torch.int32, device="cuda")
    led = torch.zeros(B, dtype=torch.float64, device="cuda")
    o = Options() if opts is None else opts
    s = Stats()
    torch.cuda.synchronize()
    if not currents:
        _check(lib().kh_batch_ac(topo.h, B, _p(params), C.byref(o), _p(f), nf, _p(nodeset), _p(x), _p(st),
                                 _p(done), _p(led), C.byref(s)))
        return Result(torch.view_as_complex(x), st, ledger=led, stats=s.as_dict(), points=done)
    cur = torch.zeros((nf, topo.n_elems, B, 2), dtype=torch.float64, device="cuda")
    lf = torch.zeros((nf, 6, B), dtype=torch.float64, device="cuda")
    out = AcOutputs(d_xac=x.data_ptr(), d_iac=cur.data_ptr(), d_ledger_f=lf.data_ptr(), d_status=st.data_ptr(),
                    d_points=done.data_ptr(), d_ledger=led.data_ptr())
    _check(lib().kh_batch_ac_ex(topo.h, B, _p(params), C.byref(o), _p(f), nf, _p(nodeset), C.byref(out),
                                C.byref(s)))
    r = Result(torch.view_as_complex(x), st, ledger=led, stats=s.as_dict(), points=done)
    r.currents = torch.view_as_complex(cur)
    r.ledgers = lf
    return r


class Group(C.Structure):
    _fields_ = [("topology", C.c_void_p), ("B", C.c_int32), ("reserved", C.c_int32), ("d_params", C.c_void_p),
                ("d_lane_spec", C.c_void_p), ("d_x", C.c_void_p), ("d_status", C.c_void_p), ("d_iters", C.c_void_p),
                ("d_accepted", C.c_void_p), ("d_rejected", C.c_void_p), ("d_ledger", C.c_void_p),
                ("d_nodeset", C.c_void_p), ("d_report", C.c_void_p), ("d_report_ledger", C.c_void_p)]


REPORT_FIELDS = ("voltage", "current", "power", "storage", "energy", "supplied", "exterior", "channel")
LEDGER_FIELDS = ("inflow", "outflow", "dissipation", "storage_rate", "model_supplied", "hypothetical", "residual",
                 "scale", "roundoff", "relative_residual")


def run_groups(topos, params_list, spec: Optional[TranSpec] = None, opts: Optional[Options] = None,
               lane_specs=None, polish: bool = False):
    """An uneven batch in one run: group g = (topos[g], params_list[g] [P_g, B_g]).
    spec None runs analysis.OP, else analysis.Tran. Returns one Result per group
    and the batch counters."""
    torch = _torch()
    n = len(topos)
    gs = (Group * n)()
    results, keep = [], []
    for g in range(n):
        prm = params_list[g]
        if isinstance(prm, np.ndarray):
            prm = to_device(prm)
        B = prm.shape[1]
        ls = None
        if spec is not None and lane_specs is not None:
            ls = lane_specs[g]
            if isinstance(ls, np.ndarray):
                ls = to_device(ls)
        r = Result(x=torch.zeros((topos[g].n, B), dtype=torch.float64, device="cuda"),
                   status=torch.zeros(B, dtype=torch.int32, device="cuda"),
                   iters=torch.zeros(B, dtype=torch.int32, device="cuda"),
                   accepted=torch.zeros(B, dtype=torch.int32, device="cuda"),
                   rejected=torch.zeros(B, dtype=torch.int32, device="cuda"),
                   ledger=torch.zeros(B, dtype=torch.float64, device="cuda"))
        keep += [prm, ls]
        gs[g].topology = topos[g].h
        gs[g].B = B
        gs[g].d_params = prm.data_ptr()
        gs[g].d_lane_spec = ls.data_ptr() if ls is not None else None
        gs[g].d_x = r.x.data_ptr()
        gs[g].d_status = r.status.data_ptr()
        gs[g].d_iters = r.iters.data_ptr()
        gs[g].d_accepted = r.accepted.data_ptr()
        gs[g].d_rejected = r.rejected.data_ptr()
        gs[g].d_ledger = r.ledger.data_ptr()
        results.append(r)
    o = Options() if opts is None else opts
    if polish:
        o.polish = 1
    s = Stats()
    torch.cuda.synchronize()
    _check(lib().kh_batch_run(n, gs, C.byref(o), C.byref(spec) if spec is not None else None, C.byref(s)))
    stats = s.as_dict()
    for r in results:
        r.stats = stats
    return results, stats


class Names(C.Structure):
    _fields_ = [("title", C.c_char_p), ("num_nodes", C.c_int32), ("n_elems", C.c_int32),
                ("node_names", C.POINTER(C.c_char_p)), ("elem_names", C.POINTER(C.c_char_p)),
                ("elems", C.c_void_p)]


FORMAT_JSON, FORMAT_CSV = 0, 1


def format_op(node_names, elem_names, elems: np.ndarray, status: int, x, report, ledger, fmt: int = FORMAT_JSON,
              title: str = "", version: Optional[str] = None, library=None) -> bytes:
    """One circuit's operating point as the reference writes it: the
    `kronspice -json` document (fmt FORMAT_JSON) or output.WriteCSV's rows
    (FORMAT_CSV). node_names [num_nodes] with "0" first, elem_names and elems
    (the element table) [n_elems]; x [n], report [n_elems, 8], ledger [10] host
    arrays (one circuit's slices of op(report=True)'s outputs)."""
    L = library if library is not None else lib()
    L.kh_format_op.restype = C.c_int64
    L.kh_format_op.argtypes = [C.c_int32, C.c_void_p, C.c_char_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_int64]
    nn = (C.c_char_p * len(node_names))(*[n.encode() for n in node_names])
    en = (C.c_char_p * max(len(elem_names), 1))(*[n.encode() for n in elem_names])
    els = np.ascontiguousarray(elems)
    nm = Names(title=title.encode(), num_nodes=len(node_names), n_elems=len(elem_names), node_names=nn,
               elem_names=en, elems=els.ctypes.data)
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (x, report, ledger)]
    ptrs = [None if a is None else a.ctypes.data for a in arrs]
    ver = None if version is None else version.encode()
    size = L.kh_format_op(fmt, C.byref(nm), ver, status, *ptrs, None, 0)
    if size < 0:
        raise KhError(BAD_ARGUMENT, "kh_format_op: invalid arguments")
    buf = C.create_string_buffer(int(size) + 1)
    L.kh_format_op(fmt, C.byref(nm), ver, status, *ptrs, buf, size)
    return buf.raw[:size]


def _names(node_names, elem_names, elems, title):
    nn = (C.c_char_p * len(node_names))(*[n.encode() for n in node_names])
    en = (C.c_char_p * max(len(elem_names), 1))(*[n.encode() for n in elem_names])
    els = np.ascontiguousarray(elems)
    nm = Names(title=title.encode(), num_nodes=len(node_names), n_elems=len(elem_names), node_names=nn,
               elem_names=en, elems=els.ctypes.data)
    return nm, (nn, en, els)


def format_dc(node_names, elem_names, elems: np.ndarray, status: int, n: int, sweeps, sources, points: int,
              sweep_x, sweep_i, max_rel: float, fmt: int = FORMAT_JSON, title: str = "",
              version: Optional[str] = None, library=None) -> bytes:
    """One circuit's DC sweep as the reference writes it: the document with a
    "dc" analysis (FORMAT_JSON) or output.WriteDCCSV's table (FORMAT_CSV).
    sweeps [(element index, start, stop, step)] inner first, sources their
    names; sweep_x [points, n] and sweep_i [points, n_elems] host arrays (one
    circuit's slices of dc(currents=True))."""
    L = library if library is not None else lib()
    L.kh_format_dc.restype = C.c_int64
    L.kh_format_dc.argtypes = [C.c_int32, C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                               C.c_int32, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int64]
    nm, keep = _names(node_names, elem_names, elems, title)
    spec = dc_spec(sweeps)
    src = (C.c_char_p * len(sources))(*[s_.encode() for s_ in sources])
    xs = None if sweep_x is None else np.ascontiguousarray(sweep_x, np.float64)
    cs = None if sweep_i is None else np.ascontiguousarray(sweep_i, np.float64)
    ver = None if version is None else version.encode()
    args = [fmt, C.byref(nm), ver, status, n, C.byref(spec), src, points, None if xs is None else xs.ctypes.data,
            None if cs is None else cs.ctypes.data, max_rel]
    size = L.kh_format_dc(*args, None, 0)
    if size < 0:
        raise KhError(BAD_ARGUMENT, "kh_format_dc: invalid arguments")
    buf = C.create_string_buffer(int(size) + 1)
    L.kh_format_dc(*args, buf, size)
    return buf.raw[:size]


def format_ac(node_names, elem_names, elems: np.ndarray, status: int, n: int, freqs, xac, iac, ledger_f,
              max_rel: float, fmt: int = FORMAT_JSON, title: str = "", version: Optional[str] = None,
              library=None) -> bytes:
    """One circuit's AC sweep as the reference writes it: the document with an
    "ac" analysis (FORMAT_JSON) or output.WriteACCSV's table (FORMAT_CSV).
    xac [nf, n] and iac [nf, n_elems] complex, ledger_f [nf, 6] host arrays (one
    circuit's slices of ac(currents=True))."""
    L = library if library is not None else lib()
    L.kh_format_ac.restype = C.c_int64
    L.kh_format_ac.argtypes = [C.c_int32, C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int64]
    nm, keep = _names(node_names, elem_names, elems, title)
    fr = np.ascontiguousarray(freqs, np.float64)
    arrs = [None if a is None else np.ascontiguousarray(a) for a in (xac, iac, ledger_f)]
    for a, t in zip(arrs, (np.complex128, np.complex128, np.float64)):
        if a is not None and a.dtype != t:
            raise ValueError(f"format_ac: want {t}, got {a.dtype}")
    ver = None if version is None else version.encode()
    args = [fmt, C.byref(nm), ver, status, n, len(fr), fr.ctypes.data] + \
        [None if a is None else a.ctypes.data for a in arrs] + [max_rel]
    size = L.kh_format_ac(*args, None, 0)
    if size < 0:
        raise KhError(BAD_ARGUMENT, "kh_format_ac: invalid arguments")
    buf = C.create_string_buffer(int(size) + 1)
    L.kh_format_ac(*args, buf, size)
    return buf.raw[:size]


class Candidate(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("n_elems", C.c_int32), ("nparams", C.c_int32), ("reserved", C.c_int32),
                ("elems", C.c_void_p), ("params", C.c_void_p), ("lane_spec", C.c_void_p)]


def run_generation(candidates, spec: Optional[TranSpec] = None, opts: Optional[Options] = None):
    """A whole generation through kh_generation_run: candidates is a list of
    (num_nodes, elements, params [P], lane_spec [4] or None) of any mix of
    topologies. Returns (x list of [n_i] arrays, status, iters, accepted,
    rejected, stats), everything at the candidate's own index."""
    L = lib()
    n = len(candidates)
    cs = (Candidate * n)()
    keep = []
    sizes = []
    for i, (nn, els, prm, lane) in enumerate(candidates):
        els = np.ascontiguousarray(els)
        prm = np.ascontiguousarray(prm, np.float64)
        lane = None if lane is None else np.ascontiguousarray(lane, np.float64)
        keep += [els, prm, lane]
        cs[i].num_nodes, cs[i].n_elems, cs[i].nparams = nn, len(els), len(prm)
        cs[i].elems = els.ctypes.data
        cs[i].params = prm.ctypes.data
        cs[i].lane_spec = lane.ctypes.data if lane is not None else None
        k = L.kh_candidate_size(C.byref(cs[i]))
        if k < 0:
            raise KhError(-k, f"candidate {i}: invalid circuit")
        sizes.append(k)
    xs = [np.zeros(k) for k in sizes]
    ptrs = (C.c_void_p * n)(*[x.ctypes.data for x in xs])
    st, it, acc, rej = (np.zeros(n, np.int32) for _ in range(4))
    o = Options() if opts is None else opts
    s = Stats()
    _check(L.kh_generation_run(n, cs, C.byref(o), C.byref(spec) if spec is not None else None, ptrs, _p(st), _p(it),
                               _p(acc), _p(rej), C.byref(s)))
    return xs, st, it, acc, rej, s.as_dict()


def tran(topo: Topology, params, spec: TranSpec, opts: Optional[Options] = None, lane_spec=None) -> Result:
    """analysis.Tran for every circuit; lane_spec [4, B] gives per-circuit step, stop, max step, min step."""
    torch = _torch()
    if isinstance(params, np.ndarray):
        params = to_device(params)
    if lane_spec is not None and isinstance(lane_spec, np.ndarray):
        lane_spec = to_device(lane_spec)
    B = params.shape[1]
    x = torch.zeros((topo.n, B), dtype=torch.float64, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    acc = torch.zeros(B, dtype=torch.int32, device="cuda")
    rej = torch.zeros(B, dtype=torch.int32, device="cuda")
    led = torch.zeros(B, dtype=torch.float64, device="cuda")
    o = Options() if opts is None else opts
    s = Stats()
    torch.cuda.synchronize()
    _check(lib().kh_batch_tran(topo.h, B, _p(params), C.byref(o), C.byref(spec), _p(lane_spec), _p(x), _p(st),
                               _p(acc), _p(rej), _p(led), C.byref(s)))
    return Result(x, st, accepted=acc, rejected=rej, ledger=led, stats=s.as_dict())


_mlp_bound = False


def _mlp_lib():
    """The library with the kh_mlp_* signatures bound."""
    global _mlp_bound
    L = lib()
    if not _mlp_bound:
        P, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.kh_mlp_create.argtypes = [i32, i32, i32, C.c_uint64, P]
        L.kh_mlp_forward.argtypes = [P, i32, P, P]
        L.kh_mlp_gradients.argtypes = [P, i32, P, P, P]
        L.kh_mlp_adam_step.argtypes = [P, C.c_float]
        L.kh_mlp_copy.argtypes = [P, i32, P, i64, i32]
        L.kh_mlp_last_error.restype = C.c_char_p
        L.kh_mlp_destroy.argtypes = [P]
        L.kh_mlp_destroy.restype = None
        _mlp_bound = True
    return L


class MLP:
    """The learned warm start's network on the GPU (kh_mlp_*): n_in -> hidden ->
    hidden -> n_out, ReLU, bf16 matrix-core products with FP32 accumulation,
    Adam on FP32 master weights."""

    def __init__(self, n_in: int, n_out: int, hidden: int = 256, seed: int = 1):
        self.n_in, self.n_out, self.hidden = n_in, n_out, hidden
        h = C.c_void_p()
        st = _mlp_lib().kh_mlp_create(n_in, hidden, n_out, seed, C.byref(h))
        if st != OK:
            raise KhError(st, _mlp_lib().kh_mlp_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.kh_mlp_destroy(self.h)
            self.h = None

    def _check(self, st):
        if st != OK:
            raise KhError(st, _mlp_lib().kh_mlp_last_error().decode())

    @property
    def count(self) -> int:
        h = self.hidden
        return self.n_in * h + h + h * h + h + h * self.n_out + self.n_out

    def forward(self, x):
        """x [B, n_in] float32 device tensor -> [B, n_out]."""
        torch = _torch()
        x = x.contiguous()
        y = torch.empty((x.shape[0], self.n_out), dtype=torch.float32, device=x.device)
        self._check(_mlp_lib().kh_mlp_forward(self.h, x.shape[0], _p(x), _p(y)))
        return y

    def gradients(self, x, t) -> float:
        """Mean squared error of the prediction against t [B, n_out]; the
        gradients are held for adam_step / parameters(grads=True)."""
        loss = C.c_float()
        x, t = x.contiguous(), t.contiguous()
        self._check(_mlp_lib().kh_mlp_gradients(self.h, x.shape[0], _p(x), _p(t), C.byref(loss)))
        return loss.value

    def adam_step(self, lr: float):
        self._check(_mlp_lib().kh_mlp_adam_step(self.h, lr))

    def parameters(self, grads: bool = False) -> np.ndarray:
        out = np.zeros(self.count, np.float32)
        self._check(_mlp_lib().kh_mlp_copy(self.h, 1 if grads else 0, _p(out), out.size, 0))
        return out

    def set_parameters(self, values: np.ndarray):
        v = np.ascontiguousarray(values, np.float32)
        self._check(_mlp_lib().kh_mlp_copy(self.h, 0, _p(v), v.size, 1))
