/*
 * kronhip.h — C ABI of the MI355X solve path for kronspice circuits.
 *
 * The library replaces the linear-solver seam of kronspice's analysis package
 * (analysis.Options.Solver / SolverFactory, analysis/op.go:50-70, and the
 * solver.Matrix interface, solver/matrix.go:23-46) and adds batched entry
 * points that run analysis.OP / analysis.Tran semantics (analysis/op.go:331,
 * analysis/tran.go:70) for many circuits that share one topology.
 *
 * All signatures use plain C types: int32/int64, double, pointers, sizes.
 * Device pointers are hipMalloc'd memory on the current device; host pointers
 * are ordinary process memory. Every call returns a kh_status; the library
 * never aborts the process.
 */
#ifndef KRONHIP_H
#define KRONHIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (per call and per circuit) ------------------------------
 * Mirror the sentinel errors of solver/matrix.go:16-20 (ErrSingularMatrix,
 * ErrBadSize, ErrNonFinite), analysis/newton.go:20 (ErrNoConvergence),
 * analysis/tran.go:18-22 (ErrTimestepTooSmall, ErrTooManySteps,
 * ErrBadTranSpec) and em/energy.go:18 (ErrLedgerResidual). */
typedef int32_t kh_status;
enum {
  KH_OK = 0,
  KH_SINGULAR = 1,          /* solver.ErrSingularMatrix */
  KH_NONFINITE = 2,         /* solver.ErrNonFinite */
  KH_BADSIZE = 3,           /* solver.ErrBadSize */
  KH_NOCONV = 4,            /* analysis.ErrNoConvergence */
  KH_TSTEP_TOO_SMALL = 5,   /* analysis.ErrTimestepTooSmall */
  KH_TOO_MANY_STEPS = 6,    /* analysis.ErrTooManySteps */
  KH_INTERRUPTED = 7,       /* context deadline (analysis/op.go:81-89) */
  KH_LEDGER = 8,            /* em.ErrLedgerResidual */
  KH_BAD_TRAN_SPEC = 9,     /* analysis.ErrBadTranSpec */
  KH_UNSUPPORTED = 10,      /* device.ErrUnsupported */
  KH_BAD_ARGUMENT = 11,     /* invalid handle, pointer or shape */
  KH_HIP_ERROR = 12,        /* a HIP runtime call failed */
  KH_NO_DEVICE = 13,        /* no usable gfx950 device */
  KH_CAPACITY = 14,         /* an L/U capacity bound was exceeded */
  KH_RUNNING = 15           /* internal: circuit still in progress */
};

/* ---- circuit description ---------------------------------------------------
 * A circuit is a node count (ground = node 0 included, as
 * circuit.Circuit.NumNodes) and an ordered element list. Node numbering is
 * the reference builder's: order of first appearance (circuit/builder.go:38-48).
 * Element order is the device order of device.FromCircuit
 * (device/device.go:347-451), which fixes the unknown numbering and the
 * stamp order (analysis/op.go:295-299).
 *
 * node[] per kind (as circuit.Element Nodes + Extra, circuit/builder.go:101-118):
 *   R C L V I D : node[0] = n+, node[1] = n-
 *   M (MOSFET)  : node[0] = drain, node[1] = source, node[2] = gate, node[3] = bulk
 *   Q (BJT)     : node[0] = collector, node[1] = emitter, node[2] = base, node[3] = substrate (0: none)
 *   E G (VCVS, VCCS): node[0] = n+, node[1] = n-, node[2] = nc+, node[3] = nc-
 *   F H (CCCS, CCVS): node[0] = n+, node[1] = n-, node[2] = index in the element list of the
 *                     controlling element (a V, L, E or H: its branch current, positive into its n+)
 *   K (mutual)      : node[0], node[1] = indices in the element list of the two coupled
 *                     inductors (circuit.Element.Couples, device/mutual.go:57-72)
 *
 * Each element reads its parameters at params[poff ...]; the per-kind layout
 * is below (already resolved by the host setup step: defaults, area factors,
 * process-parameter derivation, waveform defaults for the analysis). */
enum {
  KH_ELEM_R = 1, KH_ELEM_C = 2, KH_ELEM_L = 3, KH_ELEM_V = 4,
  KH_ELEM_I = 5, KH_ELEM_D = 6, KH_ELEM_M = 7, KH_ELEM_Q = 8,
  KH_ELEM_E = 10, KH_ELEM_F = 11, KH_ELEM_G = 12, KH_ELEM_H = 13,  /* linear controlled sources */
  KH_ELEM_K = 14                                                   /* mutual inductance */
};

typedef struct kh_element {
  int32_t kind;
  int32_t node[4];
  int32_t poff;
} kh_element;

/* Parameter layouts (doubles at params[poff + k]).
 * R: 0 R, 1 g = 1/R                                       (device/passive.go:65-69)
 * C: 0 C, 1 IC, 2 hasIC                                   (device/reactive.go:43-47)
 * L: 0 L, 1 IC, 2 hasIC                                   (device/reactive.go:214-218)
 * V, I: 0 dc, 1 hasDC, 2 wave kind (0 none, 1 PULSE, 2 SIN, 3 EXP),
 *       3..9 resolved wave args a[0..6] (device/waveform.go:25-66),
 *       10 AC magnitude, 11 AC phase [deg]                (device/source.go:21-35)
 * D: 0 IS, 1 N, 2 RS, 3 CJO, 4 VJ, 5 M, 6 TT, 7 BV, 8 IBV, 9 FC, 10 hasBV,
 *    11 vte, 12 vcrit, 13 xbv, 14 grs, 15 hasCharge       (device/diode.go:85-128)
 * M (levels 1-3): 0 typ, 1 beta, 2 vbi, 3 vto (type-normalized), 4 GAMMA,
 *    5 PHI, 6 LAMBDA, 7 gdpr, 8 gspr, 9 isd, 10 iss, 11 vcritD, 12 vcritS,
 *    13 czbd, 14 czbdsw, 15 czbs, 16 czbssw, 17 PB, 18 MJ, 19 MJSW, 20 FC,
 *    21..23 fbt[3], 24..26 fsw[3], 27 cgso, 28 cgdo, 29 cgbo, 30 oxideCap,
 *    31 vt, 32 hasJcap, 33 hasGcap                        (device/mosfet.go:218-323)
 *    34 level (1, 2, 3), 35 NFS, 36 XJ, 37 UCRIT, 38 UEXP, 39 VMAX, 40 NEFF,
 *    41 DELTA, 42 THETA, 43 ETA, 44 KAPPA, 45 UO, 46 leff, 47 weff, 48 cox,
 *    49 xd, 50 alpha, 51 cfs, 52 factor, 53 narrow, 54 eta3, 55 sqrtPB,
 *    56 hasNSUB, 57 LD        (device/mosfet.go:128-324, device/mosfet_levels.go:22-55)
 * Q (Gummel-Poon, area applied): 0 type (+1 NPN, -1 PNP), 1 IS, 2 BF, 3 NF,
 *    4 VAF, 5 IKF, 6 ISE, 7 NE, 8 BR, 9 NR, 10 VAR, 11 IKR, 12 ISC, 13 NC,
 *    14 RB, 15 IRB, 16 RBM, 17 RE, 18 RC, 19 CJE, 20 VJE, 21 MJE, 22 TF,
 *    23 XTF, 24 VTF, 25 ITF, 26 CJC, 27 VJC, 28 MJC, 29 XCJC, 30 TR, 31 CJS,
 *    32 VJS, 33 MJS, 34 FC, 35 ISS, 36 NS, 37 SUBS (+1 vertical, -1 lateral),
 *    38 vt, 39 vcrit, 40 rbpr, 41 rbpi, 42 gcpr, 43 gepr, 44 ovtf,
 *    45..47 f1e f2e f3e, 48..50 f1c f2c f3c, 51 hasCharge  (device/bjt.go:142-242)
 * E, F, G, H: 0 gain (V/V, A/A, S, ohm)                  (device/controlled.go:45-51)
 * K: 0 M = k·sqrt(L1·L2) [H]                              (device/mutual.go:44-50, device/device.go:407-416)
 */
#define KH_NPAR_R 2
#define KH_NPAR_C 3
#define KH_NPAR_L 3
#define KH_NPAR_SRC 12
#define KH_NPAR_D 16
#define KH_NPAR_M 58
#define KH_NPAR_Q 52
#define KH_NPAR_CTL 1
#define KH_NPAR_K 1

/* Analysis options: analysis.Options (analysis/op.go:22-62) with the same
 * zero-value defaults (analysis/op.go:91-110). */
typedef struct kh_options {
  double reltol;            /* 0 -> 1e-3 */
  double abstol;            /* 0 -> 1e-12 */
  double vntol;             /* 0 -> 1e-6 */
  double gmin;              /* 0 -> 1e-12 */
  int32_t itl1;             /* 0 -> 100 */
  int32_t itl4;             /* 0 -> 10 */
  int32_t no_gmin_stepping;
  int32_t no_source_stepping;
  int32_t no_pseudo_transient;
  int32_t skip_energy_check;
  double ledger_tol;        /* 0 -> 1e-9 (em.DefaultTolerance) */
  int32_t polish;           /* kh_batch_op: polish the operating point to round-off
                               after convergence (Sim.polish, analysis/op.go:243-271) */
  int32_t circuits_per_wave; /* batched analyses: circuits sharing one wavefront
                               (1..64; 0 -> 1) */
  int32_t staging;          /* batched analyses: 0 -> circuits are worked on in place in
                               device memory; 1 -> during a launch each circuit holds in LDS
                               its whole working set when it fits its share of a compute
                               unit's LDS (the share that keeps the launch resident), else
                               its most used arrays up to that share */
  int32_t waves_per_simd;   /* batched analyses: register budget of the persistent kernel,
                               as resident wavefronts per SIMD: 0 or 1 -> unconstrained,
                               2 or 4 -> fewer registers (spilled), more circuits in flight */
  int32_t backend;          /* Options.Backend (solver/matrix.go:49-93): 0 auto (dense LU below
                               30 unknowns, sparse GP LU from 30), 1 dense, 2 sparse */
  int32_t precision;        /* sparse LU arithmetic: 0 -> FP64 (the reference's); 1 -> FP32
                               factors with FP64 iterative refinement, pivot sequence certified
                               equal to FP64's, every uncertified system redone in FP64
                               (no reference counterpart: the reference factors in FP64 only) */
  int32_t circuit_wave;     /* batched analyses: 0 -> each circuit runs on one lane (k_persist,
                               circuits_per_wave / staging / waves_per_simd apply); 1 -> each
                               circuit runs on a whole wavefront, its stamping, tests and LU
                               spread over the lanes, circuits taken largest first from a queue
                               (k_wave; FP64 LU: precision is not applied). Switched off in this
                               build: 1 returns KH_UNSUPPORTED (its first device run did not
                               finish); the CPU emulation of its device code checks it */
  int32_t no_schedule;      /* AC sweeps: 0 -> the systems of a pattern that follow one probe system's
                               pivot rows are factored in lockstep along that factorization's
                               compiled structure (bit-identical; any system whose own pivot rule
                               departs from it is refactored alone); 1 -> every system alone.
                               precision 1 and the dense backend do not use schedules */
} kh_options;

/* Transient request: analysis.TranSpec + TranOptions (analysis/tran.go:26-59). */
typedef struct kh_tran_spec {
  double step, stop, start, max_step;
  int32_t uic;
  int32_t method;           /* 0 or 2 -> trapezoidal, 1 -> backward Euler */
  double chgtol;            /* 0 -> 1e-14 */
  double trtol;             /* 0 -> 7 */
  double min_step;          /* 0 -> 1e-12 * stop */
  int64_t max_steps;        /* 0 -> 1,000,000 */
} kh_tran_spec;

/* Per-batch counters filled by the batched entry points. */
typedef struct kh_stats {
  int64_t solves;            /* linear solves executed (factor + solve) */
  int64_t newton_iterations; /* sum over circuits */
  int64_t accepted_steps;    /* transient: sum over circuits */
  int64_t rejected_steps;
  int64_t analyses;          /* symbolic analyses (orderings) computed on the host */
  int64_t patterns;          /* distinct sparsity patterns seen */
  int64_t ticks;             /* device iterations launched */
  int64_t bytes_algorithmic; /* sum over executed solves of 8*(nnzA + 2n) */
  double  seconds;           /* wall time of the device loop */
  double  lu_seconds;        /* summed duration of the factor+solve launches (HIP events) */
  double  control_seconds;   /* summed duration of the stamp/control launches (HIP events) */
  int64_t lu_launches;
  int64_t control_launches;
  double  cycles_lu;         /* summed over circuits: shader clock cycles spent in factor+solve */
  double  cycles_control;    /* ... and in the rest of the analysis (stamping, Newton, LTE, ledger) */
  int64_t refine_passes;     /* precision 1: FP32 solves of the refinement, summed */
  int64_t fp64_redos;        /* precision 1: systems the mixed path could not certify (redone in FP64) */
  /* parts of cycles_control, summed over circuits: stamping (Sim.solve's Clear + Load), Newton's
     convergence test, LTE step proposals, accepting a time point, its energy-ledger record */
  double  cycles_stamp, cycles_converge, cycles_lte, cycles_accept, cycles_ledger;
  /* staging 1: LDS bytes per circuit of the staged arrays (largest over groups, last launch), and the
     groups of the last launch that stage only part of their arrays (the rest stays in device memory) */
  int64_t staged_bytes, staged_partial;
  /* AC sweeps: systems solved along a compiled schedule, and systems that tried it and were
     refactored alone because their pivot rule departed from it (or a check failed) */
  int64_t sched_solves, sched_redos;
  /* ... and the probe systems factored on the host into schedules, the lockstep rounds run, and
     the host time spent on the probes (gathering their values and factoring them) */
  int64_t sched_probes, sched_rounds;
  double  sched_host_seconds;
  double  sched_seconds;     /* summed duration of the lockstep launches (HIP events; part of lu_seconds) */
} kh_stats;

/* ---- library ---------------------------------------------------------------*/
kh_status kh_init(void);                 /* selects the current HIP device */
const char* kh_status_string(kh_status s);
const char* kh_last_error(void);         /* detail of the last failing call (thread-local) */

/* ---- per-system seam: solver.Matrix (solver/matrix.go:23-36) --------------
 * A system is stamped with Add (accumulating, ground = negative index ignored,
 * solver/sparse.go:114-143), solved with Gilbert–Peierls LU after minimum-degree
 * ordering (solver/sparse.go:154-401) on the GPU, and exposes AbsProduct
 * (solver/sparse.go:481-490). kh_sys_add_batch is the batched form of Add: the
 * Go shim stages Add calls host-side and crosses cgo once per Solve. */
typedef struct kh_sys_s* kh_sys;
kh_status kh_sys_create(int32_t n, kh_sys* out);                     /* solver.NewSparse */
kh_status kh_sys_clear(kh_sys s);                                    /* Matrix.Clear */
kh_status kh_sys_add_batch(kh_sys s, const int32_t* i, const int32_t* j,
                           const double* v, int64_t count);          /* Matrix.Add x count */
kh_status kh_sys_solve(kh_sys s, const double* b, double* x);        /* Matrix.Solve */
kh_status kh_sys_absproduct(kh_sys s, const double* x, double* y);   /* Matrix.AbsProduct */
kh_status kh_sys_nnz(kh_sys s, int64_t* a, int64_t* l, int64_t* u);  /* Sparse.NNZ */
kh_status kh_sys_pivots(kh_sys s, int32_t* q, int32_t* prow);        /* ordering q and pivot rows of the held factorization */
void kh_sys_destroy(kh_sys s);

/* kh_sys_solve without device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Same result as
 * kh_sys_solve, bit for bit, from the same factorization; the system's
 * pattern, values, right-hand side, solution and LU workspace live in a
 * 32 MiB static device array per GPU, which the process holds until it ends.
 * A system whose workspace exceeds it is refused with KH_BADSIZE. Calls on
 * one GPU share the array: calls that overlap on one GPU race. */
kh_status kh_sys_solve_static(kh_sys s, const double* b, double* x);

/* ---- complex per-system seam: solver.CMatrix (solver/matrix.go:39-46) --------
 * CSparse (solver/sparse.go:495-531) behind the CMatrix interface, which
 * analysis.AC obtains from SolverFactory.CMatrix (analysis/op.go:67-78):
 * complex128 stamps accumulated in first-Add order, Gilbert–Peierls LU with
 * pivoting on cmplx.Abs magnitudes on the GPU. Complex values are (re, im)
 * double pairs: v [2*count], b and x [2*n]. AbsProduct (solver/sparse.go:
 * 519-527) writes y [n] = sum_j |A_ij|·|x_j| with cmplx.Abs magnitudes. */
typedef struct kh_csys_s* kh_csys;
kh_status kh_csys_create(int32_t n, kh_csys* out);                   /* solver.NewCSparse */
kh_status kh_csys_clear(kh_csys s);                                  /* CMatrix.Clear */
kh_status kh_csys_add_batch(kh_csys s, const int32_t* i, const int32_t* j,
                            const double* v, int64_t count);         /* CMatrix.Add x count */
kh_status kh_csys_solve(kh_csys s, const double* b, double* x);      /* CMatrix.Solve */
kh_status kh_csys_absproduct(kh_csys s, const double* x, double* y); /* CMatrix.AbsProduct */
kh_status kh_csys_nnz(kh_csys s, int64_t* a, int64_t* l, int64_t* u);  /* CSparse.NNZ */
kh_status kh_csys_pivots(kh_csys s, int32_t* q, int32_t* prow);      /* ordering q and pivot rows */
void kh_csys_destroy(kh_csys s);

/* ---- batched numeric LU on one pattern ---------------------------------------
 * B systems share the entry list (ri[k], cj[k]), k < nnz, in first-Add order;
 * values are SoA on the device: vals[k*B + c]; b and x are [n*B] SoA
 * (x[i*B + c]). Each system is factored and solved exactly as
 * sparseLU.factor + solve (solver/sparse.go:199-401), pivoting per system.
 * status[c] receives KH_OK / KH_SINGULAR / KH_NONFINITE; prow (optional,
 * [n*B]) the pivot row of each step. */
typedef struct kh_lu_batch_s* kh_lu_batch;
kh_status kh_lu_batch_create(int32_t n, int64_t nnz, const int32_t* ri, const int32_t* cj,
                             int32_t B, kh_lu_batch* out);
kh_status kh_lu_batch_solve(kh_lu_batch h, const double* d_vals, const double* d_b,
                            double* d_x, int32_t* d_status, int32_t* d_prow, void* stream);
/* Mixed precision form of kh_lu_batch_solve (no reference counterpart: the
 * reference factors in FP64 only). Each system is factored in FP32 with the
 * same reach and pivot rule, its pivot sequence certified equal to the FP64
 * factorization's, and solved with FP64 iterative refinement; a system that
 * cannot be certified is redone exactly as kh_lu_batch_solve does.
 * d_passes (optional, [B]): FP32 solves of the refinement, 0 where the system
 * was redone in FP64. */
kh_status kh_lu_batch_solve_mixed(kh_lu_batch h, const double* d_vals, const double* d_b, double* d_x,
                                  int32_t* d_status, int32_t* d_prow, int32_t* d_passes, void* stream);
kh_status kh_lu_batch_info(kh_lu_batch h, int32_t* q, int64_t* nnzL, int64_t* nnzU,
                           int64_t* capL, int64_t* capU);
void kh_lu_batch_destroy(kh_lu_batch h);

/* ---- JacobianTearing: torn solves on the GPU (kron/jacobian.go:66-712) -------
 * kh_tear is a kron.JacobianTearing whose cut is resolved to node-unknown pairs:
 * cuts [2*ncuts], -1 = ground, as NewJacobianTearing builds them from the torn
 * elements' terminals (kron/jacobian.go:95-170; the name resolution stays with
 * the caller). reuse_tol is JacobianTearing.ReuseTol (:71-73). Its matrices
 * (kh_tmat: JacobianTearing.Matrix / CMatrix, :180-191, the SolverFactory seam
 * of analysis.Options.Solver, analysis/op.go:50-70) learn their pattern from
 * the stamps in first-Add order, as tornMat does (:199-253), and solve on the
 * GPU by tearing (:393-589): the pattern's pieces (components once the cut
 * pairs' couplings are dropped), each factored by an equilibrated dense LU and
 * kept while its block is unchanged (or moved by at most reuse_tol of its
 * largest entry), the interconnection system of the coupling rows, and
 * refinement against the whole matrix to a componentwise backward error of
 * 1e-14, accepted at 1e-12 (kron/diakoptics.go:83-85); a system refused on the
 * way (a singular piece, no convergence, a piece over solver.MaxDense) is solved
 * by the untorn LU of solver.Auto (dense below 30 unknowns, else sparse).
 * Complex values are (re, im) double pairs. kh_tear_batch solves B systems of
 * one pattern (entries ri, ci in first-Add order) in one call, one system per
 * lane, each keeping its own piece factors between calls: d_vals [nslots*B],
 * d_b and d_x [n*B] device arrays, SoA (entry k of system c at [k*B + c]);
 * d_status [B] receives KH_OK, KH_SINGULAR or KH_NONFINITE per system. The
 * counters of every matrix and batch add up in their kh_tear, as JacobianStats do (
 * :80-92). A kh_tear, its matrices and batches are used by one thread at a time. */
typedef struct kh_tear_s* kh_tear;
typedef struct kh_tear_stats {
  int64_t solves;        /* systems solved */
  int64_t factors;       /* pieces factored */
  int64_t reuses;        /* pieces whose unchanged factors were reused */
  int64_t stale_reuses;  /* pieces reused within reuse_tol */
  int64_t fallbacks;     /* systems solved by the untorn LU */
  int64_t refinements;   /* refinement steps */
  int64_t max_pieces;    /* the most pieces of one system's plan */
} kh_tear_stats;
kh_status kh_tear_create(const int32_t* cuts, int32_t ncuts, double reuse_tol, kh_tear* out);  /* NewJacobianTearing */
kh_status kh_tear_get_stats(kh_tear t, kh_tear_stats* out);                                  /* JacobianTearing.Stats */
void kh_tear_destroy(kh_tear t);

typedef struct kh_tmat_s* kh_tmat;
kh_status kh_tmat_create(kh_tear t, int32_t n, int32_t complex_values, kh_tmat* out);  /* Matrix(n) / CMatrix(n) */
kh_status kh_tmat_clear(kh_tmat m);                                                 /* tornMat.Clear */
kh_status kh_tmat_add_batch(kh_tmat m, const int32_t* i, const int32_t* j,
                            const double* v, int64_t count);                       /* tornMat.Add x count */
kh_status kh_tmat_solve(kh_tmat m, const double* b, double* x);                     /* tornMat.Solve */
kh_status kh_tmat_absproduct(kh_tmat m, const double* x, double* y);                /* tornMat.AbsProduct */
/* Pieces of the plan the last solve used (*npieces = -1 before the first), and
 * the piece of each unknown (piece_of [n], optional). */
kh_status kh_tmat_pieces(kh_tmat m, int32_t* npieces, int32_t* piece_of);
void kh_tmat_destroy(kh_tmat m);

typedef struct kh_tear_batch_s* kh_tear_batch;
kh_status kh_tear_batch_create(kh_tear t, int32_t n, int64_t nslots, const int32_t* ri, const int32_t* ci,
                               int32_t B, int32_t complex_values, kh_tear_batch* out);
kh_status kh_tear_batch_solve(kh_tear_batch h, const double* d_vals, const double* d_b, double* d_x,
                              int32_t* d_status, void* stream);
kh_status kh_tear_batch_pieces(kh_tear_batch h, int32_t* npieces, int32_t* piece_of);
void kh_tear_batch_destroy(kh_tear_batch h);

/* ---- batched analyses of circuits sharing one topology ---------------------
 * A topology is the element list (kinds, nodes, parameter offsets) with nparams
 * doubles per circuit. Parameters for B circuits are SoA on the device:
 * params[p*B + c]. All parameters that decide the unknown layout (RS > 0 of a
 * diode, RD/RS > 0 of a MOSFET) must agree across the batch. */
typedef struct kh_topology_s* kh_topology;
kh_status kh_topology_create(int32_t num_nodes, const kh_element* elems, int32_t n_elems,
                             int32_t nparams, const double* params_c0, kh_topology* out);
int32_t kh_topology_size(kh_topology t);        /* number of MNA unknowns n */
void kh_topology_destroy(kh_topology t);

/* An uneven batch: circuits of several topologies in one run (one group per
 * topology). The run uses one persistent launch for all groups; every output
 * pointer is optional device memory: d_x [n*B], the rest [B]. */
typedef struct kh_group {
  kh_topology topology;
  int32_t B;
  int32_t reserved;
  const double* d_params;     /* [nparams*B] */
  const double* d_lane_spec;  /* transient only, optional [4*B] (see kh_batch_tran) */
  double* d_x;
  int32_t* d_status;
  int32_t* d_iters;           /* Newton iterations (solves) per circuit */
  int32_t* d_accepted;
  int32_t* d_rejected;
  double* d_ledger;           /* transient: MaxRelativeResidual per circuit */
  const double* d_nodeset;    /* optional [(numNodes-1)*B]: Newton's starting node voltages
                                 per circuit, node k at row k-1 (Options.NodeSet,
                                 analysis/op.go:39-43, applied as analysis/newton.go:221-243:
                                 the other unknowns start at 0, InitJct off) */
  double* d_report;           /* optional [B][n_elems][KH_REPORT_FIELDS]: each element's entry of
                                 output.Result at the circuit's final point (output/result.go:15-35,
                                 built as analysis/op.go:361-404), fields below */
  double* d_report_ledger;    /* optional [B][KH_LEDGER_FIELDS]: the circuit's output.Ledger totals
                                 (em.BuildLedger, em/energy.go:89-119; output/result.go:58-109) */
} kh_group;

/* Fields of one element's report (d_report) and of one circuit's ledger
 * (d_report_ledger), in this order. Channel: 0 electrical, 2 thermal, -1 the
 * element has no modeled exterior port that can carry power. */
enum {
  KH_RP_VOLTAGE = 0, KH_RP_CURRENT, KH_RP_POWER, KH_RP_STORAGE, KH_RP_ENERGY, KH_RP_SUPPLIED,
  KH_RP_EXTERIOR, KH_RP_CHANNEL, KH_REPORT_FIELDS
};
enum {
  KH_LG_INFLOW = 0, KH_LG_OUTFLOW, KH_LG_DISSIPATION, KH_LG_STORAGE_RATE, KH_LG_MODEL_SUPPLIED,
  KH_LG_HYPOTHETICAL, KH_LG_RESIDUAL, KH_LG_SCALE, KH_LG_ROUNDOFF, KH_LG_RELATIVE, KH_LEDGER_FIELDS
};

/* Runs analysis.OP (spec == NULL) or analysis.Tran (spec != NULL) on every
 * circuit of every group. */
kh_status kh_batch_run(int32_t ngroups, const kh_group* groups, const kh_options* opt,
                       const kh_tran_spec* spec, kh_stats* stats);

/* ---- result formats (output package) ----------------------------------------
 * The reference's outputs written from what the batched analyses return, byte
 * for byte as its tools write them: the JSON document of `kronspice -json`
 * (output.Document, output/json.go:18-316, encoding/json with two-space
 * indentation) and output.WriteCSV's "name,value" rows (output/csv.go:15-48).
 * Host arrays in, bytes out; the circuit's names come from its netlist. */
struct kh_dc_spec;  /* below, with the DC sweep */
typedef struct kh_names {
  const char* title;                /* deck title ("" or NULL: none) */
  int32_t num_nodes;                /* node names [num_nodes], index 0 = ground ("0") */
  int32_t n_elems;                  /* element names [n_elems], as written in the netlist */
  const char* const* node_names;
  const char* const* elem_names;
  const kh_element* elems;          /* the element table (kinds, nodes) of the topology */
} kh_names;
enum { KH_FORMAT_JSON = 0, KH_FORMAT_CSV = 1 };
/* One circuit's operating point: the document with one "op" analysis (OPDocument,
 * output/json.go:164-200) or output.WriteCSV's rows. status is the circuit's;
 * any other than KH_OK writes the failed run's document (ok false, the error's
 * class as cmd/kronspice/main.go:284-305 names it and the status text, no
 * analyses) or no CSV. x [n], report [n_elems * KH_REPORT_FIELDS] and ledger
 * [KH_LEDGER_FIELDS] are one circuit's slices of kh_batch_run's d_x (a column),
 * d_report and d_report_ledger, copied to the host. version NULL: "dev". Writes
 * at most cap bytes to buf and returns the document's length in bytes (size
 * buf with a first call of cap 0), or -1 on invalid arguments. */
int64_t kh_format_op(int32_t format, const kh_names* names, const char* version, kh_status status, const double* x,
                     const double* report, const double* ledger, char* buf, int64_t cap);
/* One circuit's DC sweep: the document with one "dc" analysis (DCDocument,
 * output/json.go:202-221) or output.WriteDCCSV's table (output/ac.go:125-165).
 * n: the circuit's unknowns; sources [nsweeps]: the swept sources' names as the
 * .dc line gives them; points: the points solved; sweep_x [points*n] and
 * sweep_i [points*n_elems]: one circuit's slices of kh_batch_dc_ex's d_sweep_x
 * and d_sweep_i, point-major; max_rel: its d_ledger value. Return value and
 * buffer as kh_format_op. */
/* One circuit's AC sweep: the document with one "ac" analysis (ACDocument,
 * output/json.go:223-253) or output.WriteACCSV's table (output/ac.go:67-100).
 * freqs [nf]; xac [nf*n*2], iac [nf*n_elems*2] and ledger_f [nf*6] are one
 * circuit's slices of kh_batch_ac_ex's d_xac, d_iac and d_ledger_f
 * (frequency-major); max_rel its d_ledger value. Phases are atan2 of the C
 * library, within an ulp of Go's math.Atan2. Return value and buffer as
 * kh_format_op. */
int64_t kh_format_ac(int32_t format, const kh_names* names, const char* version, kh_status status, int32_t n, int32_t nf,
                     const double* freqs, const double* xac, const double* iac, const double* ledger_f,
                     double max_rel, char* buf, int64_t cap);
int64_t kh_format_dc(int32_t format, const kh_names* names, const char* version, kh_status status, int32_t n,
                     const struct kh_dc_spec* spec, const char* const* sources, int32_t points, const double* sweep_x,
                     const double* sweep_i, double max_rel, char* buf, int64_t cap);

/* ---- a whole generation (the population evaluation's batched caller) ------
 * The GA evaluates every candidate of a generation with analysis.OP or Tran on
 * a pool of goroutines, results stored by candidate index (evo/engine.go:519-612,
 * evo/fitness.go:98-190). kh_generation_run takes the generation as it comes —
 * candidates of any mix of topologies, host arrays — groups the candidates
 * that share a topology (node count, element table, parameter count, MOSFET
 * levels and whether a per-circuit time scale is given), runs every group in
 * one uneven batch (kh_batch_run) and writes each candidate's results at its
 * own index, so selection stays deterministic whatever the grouping. */
typedef struct kh_candidate {
  int32_t num_nodes;
  int32_t n_elems;
  int32_t nparams;
  int32_t reserved;
  const kh_element* elems;    /* host [n_elems] */
  const double* params;       /* host [nparams] */
  const double* lane_spec;    /* transient, optional host [4]: step, stop, max step, min step */
} kh_candidate;
/* MNA unknowns of a candidate's circuit (the length of its x_out array), or a
 * negative status when its element table is invalid. */
int32_t kh_candidate_size(const kh_candidate* c);
/* x_out (optional): per candidate a host array of kh_candidate_size doubles
 * (or null to skip it); status, iters, accepted, rejected: optional host [n]. */
kh_status kh_generation_run(int32_t n, const kh_candidate* cands, const kh_options* opt, const kh_tran_spec* spec,
                            double* const* x_out, int32_t* status, int32_t* iters, int32_t* accepted,
                            int32_t* rejected, kh_stats* stats);

/* DC sweep request: analysis.DC with one or two swept sources
 * (analysis/dc.go:22-28 DCSweep, :51-130 DC). Sweep 0 is the inner loop;
 * its points are start + k*step, k = 0 .. floor((stop-start)/step + 1e-9)
 * (DCSweep.values, :31-46); element[] are indices of V or I sources in the
 * topology's element list (the source named by DCSweep.Source). */
typedef struct kh_dc_spec {
  int32_t nsweeps;          /* 1 or 2 */
  int32_t element[2];
  int32_t reserved;
  double start[2], stop[2], step[2];
} kh_dc_spec;

/* analysis.DC for B circuits of one topology: at every point the swept
 * sources' DC values are overridden (SetDCOverride, device/source.go:179-203)
 * and analysis.OP runs on the same circuit state, so every point after the
 * first starts warm from the previous one; the sweep stops at a circuit's
 * first failing point with that point's status. d_sweep_x [points*n*B]
 * (point-major, then unknown, then circuit); d_points [B] points solved;
 * d_iters (optional) [points*B] Newton iterations per point; d_ledger
 * (optional) [B] largest ledger residual over the points; d_nodeset
 * (optional) [(numNodes-1)*B] the first point's Newton start. Replaces the
 * reference's per-circuit DC (analysis/dc.go:51). */
kh_status kh_batch_dc(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                      const kh_dc_spec* spec, const double* d_nodeset, double* d_sweep_x, int32_t* d_status,
                      int32_t* d_points, int32_t* d_iters, double* d_ledger, kh_stats* stats);

/* kh_batch_dc with the sweep's element currents too (DCSweepResult.I,
 * analysis/dc.go:87-128): d_sweep_i (optional) [points*n_elems*B], point-major,
 * then element, then circuit: each element's current (n+ -> n-) at each
 * point, after that point's Accept. The other outputs are kh_batch_dc's. */
typedef struct kh_dc_outputs {
  double* d_sweep_x;   /* [points*n*B] */
  double* d_sweep_i;   /* optional [points*n_elems*B] */
  int32_t* d_status;   /* [B] */
  int32_t* d_points;   /* optional [B] */
  int32_t* d_iters;    /* optional [points*B] */
  double* d_ledger;    /* optional [B] */
} kh_dc_outputs;
kh_status kh_batch_dc_ex(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                         const kh_dc_spec* spec, const double* d_nodeset, const kh_dc_outputs* out,
                         kh_stats* stats);

/* analysis.AC for B circuits of one topology (analysis/ac.go:76-193): the
 * operating point and polish, then at every frequency the devices' LoadAC
 * stamps, a complex Gilbert–Peierls solve (CSparse, solver/sparse.go:495-531:
 * pivoting on cmplx.Abs) and the complex-Tellegen ledger check. freqs [nf] is
 * host memory (ACSpec.Frequencies' output); d_xac [nf*n*B*2] receives the
 * solutions, frequency-major, then unknown, then circuit, each a (re, im)
 * pair; d_points [B] (optional) the frequencies solved before a circuit's
 * first failure; d_ledger [B] (optional) the largest ledger residual;
 * d_nodeset as kh_batch_dc. */
kh_status kh_batch_ac(kh_topology t, int32_t B, const double* d_params, const kh_options* opt, const double* freqs,
                      int32_t nf, const double* d_nodeset, double* d_xac, int32_t* d_status, int32_t* d_points,
                      double* d_ledger, kh_stats* stats);

/* kh_batch_ac with ACResult's element currents and per-frequency ledger too
 * (analysis/ac.go:118-173): d_iac (optional) [nf*n_elems*B*2], frequency-major,
 * then element, then circuit, (re, im): each element's current phasor (n+ ->
 * n-); d_ledger_f (optional) [nf*6*B]: per frequency and circuit the average
 * power inflow, outflow, dissipation, reactive power, model-supplied power
 * and residual (output.ACLedger). The other outputs are kh_batch_ac's. */
typedef struct kh_ac_outputs {
  double* d_xac;       /* [nf*n*B*2] */
  double* d_iac;       /* optional [nf*n_elems*B*2] */
  double* d_ledger_f;  /* optional [nf*6*B] */
  int32_t* d_status;   /* [B] */
  int32_t* d_points;   /* optional [B] */
  double* d_ledger;    /* optional [B]: MaxRelativeResidual */
} kh_ac_outputs;
kh_status kh_batch_ac_ex(kh_topology t, int32_t B, const double* d_params, const kh_options* opt, const double* freqs,
                         int32_t nf, const double* d_nodeset, const kh_ac_outputs* out, kh_stats* stats);

/* analysis.OP for B circuits (analysis/op.go:331-357): dcSolve with the
 * gmin / source / pseudo-transient fallbacks (analysis/newton.go:95-311),
 * Accept, ledger check. x_out is [n*B] SoA on the device. iters (optional) the
 * Newton iterations per circuit. */
kh_status kh_batch_op(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                      double* d_x, int32_t* d_status, int32_t* d_iters, kh_stats* stats);

/* analysis.Tran for B circuits (analysis/tran.go:70-231). Records the final
 * accepted solution per circuit (d_x, [n*B]) and the accepted / rejected step
 * counts; d_ledger (optional, [B]) the run's MaxRelativeResidual.
 * d_lane_spec (optional, device [4*B]: step, stop, max_step, min_step per
 * circuit) gives each circuit its own time scale — a batch of candidates
 * whose source frequencies differ by decades — as if each ran its own
 * analysis.Tran with that TranSpec; the other fields come from spec. */
kh_status kh_batch_tran(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                        const kh_tran_spec* spec, const double* d_lane_spec, double* d_x,
                        int32_t* d_status, int32_t* d_accepted, int32_t* d_rejected,
                        double* d_ledger, kh_stats* stats);

/* ---- learned Newton warm start --------------------------------------------
 * No counterpart in the reference: the small per-topology network that
 * predicts a circuit's DC node voltages from its value genes; its predictions
 * are Options.NodeSet guesses (kh_group.d_nodeset), and Newton with its
 * fallbacks decides the answer. Layers in -> hidden -> hidden -> out with ReLU;
 * bf16 weights and activations into FP32-accumulating matrix-core products,
 * FP32 master weights, biases and Adam moments. Inputs d_x [B][n_in] and
 * targets d_t [B][n_out] are FP32 on the device. */
typedef struct kh_mlp_s* kh_mlp;
kh_status kh_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, kh_mlp* out);
kh_status kh_mlp_forward(kh_mlp m, int32_t B, const float* d_x, float* d_y);          /* d_y [B][n_out] */
kh_status kh_mlp_gradients(kh_mlp m, int32_t B, const float* d_x, const float* d_t,
                           float* loss);    /* mean squared error and its gradients (held) */
kh_status kh_mlp_adam_step(kh_mlp m, float lr);                                       /* apply the held gradients */
/* Parameters (which = 0) or held gradients (1) as one host array in the order
 * W1 [n_in][hidden], b1, W2 [hidden][hidden], b2, W3 [hidden][n_out], b3;
 * to_device = 1 writes parameters instead of reading them. */
kh_status kh_mlp_copy(kh_mlp m, int32_t which, float* host, int64_t count, int32_t to_device);
const char* kh_mlp_last_error(void);
void kh_mlp_destroy(kh_mlp m);

#ifdef __cplusplus
}
#endif
#endif /* KRONHIP_H */
