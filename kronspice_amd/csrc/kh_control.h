// kh_control.h — the per-circuit analysis loop as a resumable state machine.
//
// kronspice runs Newton (analysis/newton.go:36-82) inside the DC fallback
// chain (dcSolve, gminStepping, sourceStepping, pseudoTransient,
// analysis/newton.go:95-311), inside OP (analysis/op.go:340-357) or the
// transient loop (analysis/tran.go:95-231). Every step of that control flow
// that is not a linear solve is restated here for one lane; the lane stops
// whenever it needs a solve (after stamping the system, Sim.solve,
// analysis/op.go:289-322) and resumes on the next launch with the solve's
// outcome. Lanes therefore advance independently — per-circuit Newton
// iteration counts, fallbacks, step rejections and time steps — while the
// host only relaunches.
#pragma once
#include "kh_devices.h"

namespace kh {

// Integer control words ([NCI] per circuit).
enum {
  CI_PHASE = 0, CI_CALLER, CI_NT_IT, CI_NT_MAX, CI_INITJCT, CI_WARM, CI_SRCSTEP, CI_STATUS, CI_ITERS,
  CI_FIRST_ERR, CI_GM_OK, CI_GM_K, CI_SS_K, CI_PT_K, CI_PT_QUIET, CI_TR_ORDER, CI_TR_HIT, CI_ACCEPTED,
  CI_REJECTED, CI_WAIT, CI_NEED_SOLVE, CI_MODE, CI_METHOD, CI_LU, CI_DC_RET, CI_PL_K, CI_PID, CI_SOLVES,
  CI_SW_I, CI_SW_J, CI_SW_IT0, CI_SW_DONE, NCI
};
// CI_WAIT: why a circuit waits for the host between launches.
enum { WAIT_NONE = 0, WAIT_PATTERN = 1, WAIT_CAPACITY = 2 };
// Double control words ([NCD] per circuit).
enum {
  CD_TIME = 0, CD_H, CD_GMIN, CD_DIAG_GMIN, CD_PTG, CD_SRCFACT, CD_GM_G, CD_GM_FACTOR, CD_SS_DONE, CD_SS_STEP,
  CD_SS_F, CD_PT_H, CD_TR_T, CD_TR_H, CD_TR_TN, CD_PEAK, CD_WORST, CD_MAXREL, CD_LEDGER, CD_MINSTEP_SEEN,
  CD_MAXSTEP_SEEN, CD_CYC_LU, CD_CYC_CTL, CD_REFINE, CD_REDO, CD_CYC_STAMP, CD_CYC_CONV, CD_CYC_LTE,
  CD_CYC_LEDGER, CD_CYC_ACCEPT, NCD
};

enum { PH_INIT = 0, PH_RUN = 1, PH_DONE = 2 };
enum { AN_OP = 0, AN_TRAN = 1, AN_OP_POLISH = 2, AN_DC = 3 };
enum {
  CALLER_DC_PLAIN = 1, CALLER_GMIN, CALLER_SRC0, CALLER_SRC, CALLER_PT, CALLER_PT_FINAL, CALLER_UIC, CALLER_TRAN,
  CALLER_POLISH
};
enum { DCRET_OP = 0, DCRET_TRAN = 1 };
enum {
  ST_OK = 0, ST_SINGULAR = 1, ST_NONFINITE = 2, ST_BADSIZE = 3, ST_NOCONV = 4, ST_TSTEP = 5, ST_MAXSTEPS = 6,
  ST_LEDGER = 8, ST_CAPACITY = 14, ST_RUNNING = 15
};

// Analysis options and transient spec, resolved on the host (defaults of
// analysis/op.go:91-110 and analysis/tran.go:42-59).
struct CtlParams {
  double relTol, absTol, vnTol, gmin, ledgerTol;
  int32_t itl1, itl4, noGmin, noSource, noPT, skipEnergy, analysis;
  int32_t precision;  // sparse LU: 0 FP64, 1 FP32 factors + FP64 refinement (kh_lu.h)
  // transient
  double step, stop, start, maxStep, chgTol, trTol, minStep;
  int64_t maxSteps;
  int32_t uic, method;
  // DC sweep (analysis/dc.go:22-46): sweep 0 is the inner loop; point k of
  // sweep s is swStart[s] + k * swStep[s], k < swCount[s]
  int32_t swN, swElem[2], swCount[2];
  double swStart[2], swStep[2];
};

struct Circ {
  int32_t n, nel, nslots, numNodes;
  const int32_t* et;        // [nel][ET_W]
  const int32_t* isV;       // [n]
  const int32_t* slot_i;    // [nslots] row of each slot
  const int32_t* slot_j;    // [nslots] column
  const int32_t* diag_slot; // [n] slot of (i, i)
  int32_t nonlin;           // any nonlinear device
  int32_t dense;            // solved by the dense backend (lu_dense): no sparsity patterns
  int32_t n_nl, n_rx, n_src;
  const int32_t* nl;        // [n_nl] nonlinear devices (D, M, Q), element order
  const int32_t* rx;        // [n_rx] devices with charge or flux (C, L, D, M, Q)
  const int32_t* src;       // [n_src] sources (V, I)
  const int32_t* order;     // [nslots] slot order of AbsProduct: row-major for the dense
                            // backend (solver/dense.go:63-73), null: slot order
  // The wavefront form (kh_wave.h): elements stamp in chunks of consecutive
  // elements; chunk c holds elements [ch_el[c], ch_el[c+1]) and touches the
  // targets tgt[ch_t[c] .. ch_t[c+1]) (a value slot, or nslots + row), target t
  // by the elements tel[tel_ptr[t] .. tel_ptr[t+1]) in element order; element k
  // records at rec_off[k] of its chunk's record. ro_ptr / ro_slot: the slots of
  // each row in AbsProduct's accumulation order.
  int32_t nchunks, rec_cap;
  const int32_t* ch_el;
  const int32_t* ch_t;
  const int32_t* tgt;
  const int32_t* tel_ptr;
  const int32_t* tel;
  const int32_t* rec_off;
  const int32_t* ro_ptr;
  const int32_t* ro_slot;
};

struct CtlArgs {
  Circ ckt;
  CtlParams p;
  int64_t B;
  const double* par;
  double* st;
  double* val;
  uint8_t* present;
  double* rhs;
  double* x;
  double* xold;
  double* xsave;    // gmin / source stepping save vector
  double* ptprev;   // pseudo-transient previous state
  double* xacc;     // transient: last accepted solution
  double* absbuf;   // AbsProduct scratch
  int32_t* ci;      // [NCI][B]
  double* cd;       // [NCD][B]
  const double* lane_spec;   // optional [4][B]: per-circuit step, stop, max step, min step
  const double* nodeset;     // optional [numNodes-1][B]: per-circuit Newton start (Options.NodeSet)
  double* sw_x;              // DC sweep: per circuit [points][n] solutions
  int32_t* sw_it;            // DC sweep: per circuit [points] Newton iterations
  double* sw_i;              // DC sweep (optional): per circuit [points][nel] element currents
  const int32_t* lu_status;  // [B] result of the last factor+solve
  int32_t* counters;         // [0] lanes still running, [1] lanes waiting for a new pattern
};

KH_HD inline bool retryable(int e) { return e == ST_NOCONV || e == ST_NONFINITE || e == ST_SINGULAR; }

// The state machine's device dispatch depends on the element kinds compiled
// in (KH_KINDS), so each kind set gets its own namespace.
// Shader clock for the per-phase cycle counters (kh_stats); 0 where the
// including unit provides none.
#ifndef KH_CLOCK
#define KH_CLOCK() ((uint64_t)0)
#endif

inline namespace KH_CAT(kinds_, KH_KINDS) {

// Loops whose iterations are independent issue the loads of kChunk
// iterations before using any of them: one lane's program is a chain of
// dependent memory latencies, and a loop that waits for each iteration's
// loads pays one latency per iteration. Arithmetic order is unchanged.
constexpr int kChunk = 8;

KH_HD inline void copy_n(double* __restrict__ d, const double* __restrict__ s, int n) {
  int i = 0;
  for (; i + kChunk <= n; i += kChunk) {
    double v[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; u++) v[u] = s[i + u];
#pragma unroll
    for (int u = 0; u < kChunk; u++) d[i + u] = v[u];
  }
  for (; i < n; i++) d[i] = s[i];
}

// One circuit's state machine over its own (unit-stride) arrays; the control
// words are worked on in place, so they live wherever the circuit's arrays do
// (device memory, or LDS when the launch stages the circuit).
//
// D is the executing form (CRTP): every loop over unknowns, elements or value
// slots is a member that D may replace — Ctl runs them on one lane, the
// wavefront form (kh_wave.h) spreads them over the lanes of a wavefront that
// all run the same circuit's state machine in step.
template <class D>
struct CtlT {
  const CtlArgs A;  // a copy: one indirection less per array access (the object lives in private memory)
  int n;
  int32_t* ci;  // [NCI]
  double* cd;   // [NCD]
  Lane L;
  CtlParams P;  // options and transient spec as seen by this circuit

  KH_HD D& self() { return *static_cast<D*>(this); }

  KH_HD CtlT(const CtlArgs& a) : A(a), n(a.ckt.n), ci(a.ci), cd(a.cd) {
    L.n = n;
    L.et = a.ckt.et; L.par = a.par; L.st = a.st; L.val = a.val; L.present = a.present; L.rhs = a.rhs;
    L.X = a.xold;
    L.grew = false;
    L.rslot = nullptr;
    L.rval = nullptr;
    L.rn = 0;
    L.rcap = 0;
    L.nslots = a.ckt.nslots;
    P = a.p;
    if (a.lane_spec) {  // per-circuit time scale: step, stop, max step, min step
      P.step = a.lane_spec[0];
      P.stop = a.lane_spec[1];
      P.maxStep = a.lane_spec[2];
      P.minStep = a.lane_spec[3];
    }
  }
  KH_HD double& X(double* v, int i) { return v[i]; }
  KH_HD const int32_t* el(int k) const { return A.ckt.et + (int64_t)k * ET_W; }

  KH_HD void finish(int status) {
    ci[CI_PHASE] = PH_DONE;
    ci[CI_STATUS] = status;
    ci[CI_NEED_SOLVE] = 0;
  }

  // ---- device dispatch -----------------------------------------------------
  KH_HD void load_all() {
    // A local copy of the device view: its fields (array bases, the load
    // context, the pattern-growth flag) stay in registers while the elements
    // stamp, instead of being read and written through this object.
    Lane lane = L;
    lane.mode = ci[CI_MODE];
    lane.method = ci[CI_METHOD];
    lane.time = cd[CD_TIME];
    lane.H = cd[CD_H];
    lane.gmin = cd[CD_GMIN];
    lane.srcFact = cd[CD_SRCFACT];
    lane.initJct = ci[CI_INITJCT] != 0;
    lane.srcStepping = ci[CI_SRCSTEP] != 0;
    const int nel = A.ckt.nel;
    const int32_t* et = A.ckt.et;
    for (int k = 0; k < nel; k++) {
      const int32_t* e = et + (int64_t)k * ET_W;
      switch (e[ET_KIND]) {
        case K_R: load_R(lane, e); break;
        case K_C: load_C(lane, e); break;
        case K_L: load_L(lane, e); break;
        case K_V: load_V(lane, e); break;
        case K_I: load_I(lane, e); break;
        case K_D: if (KH_HAS(K_D)) load_D(lane, e); break;
        case K_M: if (KH_HAS(K_M)) load_M(lane, e); break;
        case K_Q: if (KH_HAS(K_Q)) load_Q(lane, e); break;
        case K_E: if (KH_HAS(K_E)) load_ctl(lane, e); break;
        case K_F: if (KH_HAS(K_F)) load_ctl(lane, e); break;
        case K_G: if (KH_HAS(K_G)) load_ctl(lane, e); break;
        case K_H: if (KH_HAS(K_H)) load_ctl(lane, e); break;
        case K_K: if (KH_HAS(K_K)) load_K(lane, e); break;
      }
    }
    L = lane;
  }
  KH_HD OpPt operating(const Lane& L, const int32_t* e, const double* x) const { return operating_point(L, e, x); }
  KH_HD void accept(double t, int mode) {  // Sim.accept, analysis/op.go:324-328
    const Lane L = this->L;  // register copy of the device view (see load_all)
    const int nel_ = A.ckt.nel;
    const int32_t* const et_ = A.ckt.et;
    double* const x_ = A.x;
    (void)nel_;
    for (int q = 0; q < A.ckt.n_rx; q++) {
      const int32_t* e = et_ + (int64_t)A.ckt.rx[q] * ET_W;
      switch (e[ET_KIND]) {
        case K_C: accept_C(L, e, x_, t, mode); break;
        case K_L: accept_L(L, e, x_, t, mode); break;
        case K_D: if (KH_HAS(K_D)) accept_D(L, e, x_, t, mode); break;
        case K_M: if (KH_HAS(K_M)) accept_M(L, e, x_, t, mode); break;
        case K_Q: if (KH_HAS(K_Q)) accept_Q(L, e, x_, t, mode); break;
        case K_K: if (KH_HAS(K_K)) accept_K(L, e, x_); break;
      }
    }
  }
  KH_HD bool converged() {  // analysis/newton.go:59-82
    // Every test is free of side effects, so the result is the conjunction of
    // the reference's sequence of tests: no device limited, every unknown
    // within tolerance, every nonlinear device's convergence test. The unknown
    // test runs without early exits (independent loads overlap); the device
    // loops visit only the nonlinear devices.
    const Lane L = this->L;  // register copy of the device view (see load_all)
    const int32_t* const et_ = A.ckt.et;
    const int32_t* const nl = A.ckt.nl;
    const int n_nl = A.ckt.n_nl;
    double* const x_ = A.x;
    for (int q = 0; q < n_nl; q++) {
      const int32_t* e = et_ + (int64_t)nl[q] * ET_W;
      if (KH_HAS(K_D) && e[ET_KIND] == K_D && L.S(e, 6) != 0) return false;
      if (KH_HAS(K_M) && e[ET_KIND] == K_M && L.S(e, MS_LIM) != 0) return false;
      if (KH_HAS(K_Q) && e[ET_KIND] == K_Q && L.S(e, QS_LIM) != 0) return false;
    }
    const double* const xo = A.xold;
    const int32_t* const isV = A.ckt.isV;
    const double vnTol = P.vnTol, absTol = P.absTol, relTol = P.relTol;
    bool ok = true;
    int i = 0;
    for (; i + kChunk <= n; i += kChunk) {
      double xv[kChunk], old[kChunk];
      int32_t v[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; u++) {
        xv[u] = x_[i + u];
        old[u] = xo[i + u];
        v[u] = isV[i + u];
      }
#pragma unroll
      for (int u = 0; u < kChunk; u++) {
        double tol = v[u] ? vnTol : absTol;
        tol += relTol * fmax(fabs(xv[u]), fabs(old[u]));
        ok = ok && !(fabs(xv[u] - old[u]) > tol);
      }
    }
    for (; i < n; i++) {
      double xv = x_[i], old = xo[i];
      double tol = isV[i] ? vnTol : absTol;
      tol += relTol * fmax(fabs(xv), fabs(old));
      ok = ok && !(fabs(xv - old) > tol);
    }
    if (!ok) return false;
    for (int q = 0; q < n_nl; q++) {
      const int32_t* e = et_ + (int64_t)nl[q] * ET_W;
      if (KH_HAS(K_D) && e[ET_KIND] == K_D && !conv_D(L, e, x_, relTol, absTol)) return false;
      if (KH_HAS(K_M) && e[ET_KIND] == K_M && !conv_M(L, e, x_, relTol, absTol)) return false;
      if (KH_HAS(K_Q) && e[ET_KIND] == K_Q && !conv_Q(L, e, x_, relTol, absTol)) return false;
    }
    return true;
  }
  KH_HD double next_breakpoint(double t) {  // analysis/tran.go:233-243
    const Lane L = this->L;  // register copy of the device view (see load_all)
    const int nel_ = A.ckt.nel;
    const int32_t* const et_ = A.ckt.et;
    double* const x_ = A.x;
    double nb = kInf;
    (void)nel_; (void)x_;
    for (int q = 0; q < A.ckt.n_src; q++) {
      const int32_t* e = et_ + (int64_t)A.ckt.src[q] * ET_W;
      {
        if ((int)L.P(e, 2) == 0) continue;
        double b = wave_next_breakpoint(L, e, t);
        if (b < nb) nb = b;
      }
    }
    return nb;
  }
  KH_HD double truncate(double tn, double h, int order) {
    const Lane L = this->L;  // register copy of the device view (see load_all)
    const int nel_ = A.ckt.nel;
    const int32_t* const et_ = A.ckt.et;
    double* const x_ = A.x;
    TruncCtx tc;
    tc.time = tn; tc.H = h; tc.method = order;
    tc.relTol = P.relTol; tc.absTol = P.absTol; tc.vnTol = P.vnTol; tc.chgTol = P.chgTol; tc.trTol = P.trTol;
    double prop = kInf;
    (void)nel_;
    for (int q = 0; q < A.ckt.n_rx; q++) {
      const int32_t* e = et_ + (int64_t)A.ckt.rx[q] * ET_W;
      double pk = kInf;
      switch (e[ET_KIND]) {
        case K_C: pk = trunc_C(L, e, x_, tc); break;
        case K_L: pk = trunc_L(L, e, x_, tc); break;
        case K_D: if (KH_HAS(K_D)) pk = trunc_D(L, e, x_, tc); break;
        case K_M: if (KH_HAS(K_M)) pk = trunc_M(L, e, x_, tc); break;
        case K_Q: if (KH_HAS(K_Q)) pk = trunc_Q(L, e, x_, tc); break;
        default: continue;
      }
      if (pk < prop) prop = pk;
    }
    return prop;
  }
  // roundoff, analysis/op.go:275-282: AbsProduct (solver/sparse.go:481-490)
  // accumulated over the stamped slots in slot order, then the weighted sum.
  KH_HD double roundoff() {
    double* const ab = A.absbuf;
    const double* const x = A.x;
    const double* const val = A.val;
    const uint8_t* const present = A.present;
    const int32_t* const si = A.ckt.slot_i;
    const int32_t* const sj = A.ckt.slot_j;
    const int32_t* const order = A.ckt.order;
    const int ns = A.ckt.nslots;
    for (int i = 0; i < n; i++) ab[i] = 0;
    int k = 0;
    for (; k + kChunk <= ns; k += kChunk) {
      int32_t sl[kChunk], row[kChunk];
      double term[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; u++) sl[u] = order ? order[k + u] : k + u;
#pragma unroll
      for (int u = 0; u < kChunk; u++) {
        row[u] = present[sl[u]] ? si[sl[u]] : -1;
        term[u] = fabs(val[sl[u]]) * fabs(x[sj[sl[u]]]);
      }
#pragma unroll
      for (int u = 0; u < kChunk; u++)  // accumulated in slot order
        if (row[u] >= 0) ab[row[u]] += term[u];
    }
    for (; k < ns; k++) {
      const int s = order ? order[k] : k;
      if (!present[s]) continue;
      ab[si[s]] += fabs(val[s]) * fabs(x[sj[s]]);
    }
    const double* const rhs = A.rhs;
    double sum = 0;
    int i = 0;
    for (; i + kChunk <= n; i += kChunk) {
      double term[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; u++) term[u] = fabs(x[i + u]) * (ab[i + u] + fabs(rhs[i + u]));
#pragma unroll
      for (int u = 0; u < kChunk; u++) sum += term[u];
    }
    for (; i < n; i++) sum += fabs(x[i]) * (ab[i] + fabs(rhs[i]));
    return 64 * 0x1p-52 * sum;
  }
  // Ledger of an operating point: BuildLedger + RelativeResidual
  // (em/energy.go:90-119, output/result.go:100-109).
  KH_HD double ledger_op() {
    const Lane L = this->L;  // register copy of the device view (see load_all)
    const int nel_ = A.ckt.nel;
    const int32_t* const et_ = A.ckt.et;
    double* const x_ = A.x;
    double inflow = 0, outflow = 0, diss = 0, stor = 0, sup = 0, scale = 0;
    for (int k = 0; k < nel_; k++) {
      const int32_t* e = et_ + (int64_t)k * ET_W;
      OpPt o = operating(L, e, x_);
      double pw = 0;
      int ch = -1;
      int kind = e[ET_KIND];
      if (kind == K_R) {
        if (L.P(e, 0) > 0) { pw = -o.power; ch = 2; }
      } else if (kind == K_V || kind == K_I) {
        pw = -o.power; ch = 0;
      } else if (kind == K_D || kind == K_M || kind == K_Q) {
        pw = -(o.power - o.storage); ch = 2;
      }
      if (ch >= 0 && pw != 0) {
        scale += fabs(pw);
        if (ch == 2 && pw < 0) diss += -pw;
        else if (pw > 0) inflow += pw;
        else outflow += -pw;
      }
      if (kind == K_C || kind == K_L || kind == K_D || kind == K_M || kind == K_Q) {
        stor += o.storage;
        scale += fabs(o.storage);
      }
      if ((kind == K_R && L.P(e, 0) < 0) || is_controlled(kind)) {  // model-supplied power
        double ms = -o.power;
        if (ms != 0) { sup += ms; scale += fabs(ms); }
      }
    }
    double res = inflow + sup - outflow - diss - stor;
    double ro = self().roundoff();
    if (scale == 0 || fabs(res) <= fmax(1e-21, ro)) return 0;
    return fabs(res / scale);
  }
  // recorder.record's per-point check (analysis/tran.go:291-347).
  KH_HD int record(double t, bool imposed) {
    const Lane L = this->L;  // register copy of the device view (see load_all)
    const int nel_ = A.ckt.nel;
    const int32_t* const et_ = A.ckt.et;
    double* const x_ = A.x;
    double inflow = 0, outflow = 0, diss = 0, stor = 0, sup = 0, scale = 0;
    // ports of every element, in element order
    for (int k = 0; k < nel_; k++) {
      const int32_t* e = et_ + (int64_t)k * ET_W;
      int kind = e[ET_KIND];
      double pw = 0;
      int ch = -1;
      if (kind == K_R || kind == K_V || kind == K_I || kind == K_D || kind == K_M || kind == K_Q) {
        OpPt o = operating(L, e, x_);
        if (kind == K_R) { if (L.P(e, 0) > 0) { pw = -o.power; ch = 2; } }
        else if (kind == K_V || kind == K_I) { pw = -o.power; ch = 0; }
        else { pw = -(o.power - o.storage); ch = 2; }
      }
      if (ch >= 0 && pw != 0) {
        scale += fabs(pw);
        if (ch == 2 && pw < 0) diss += -pw;
        else if (pw > 0) inflow += pw;
        else outflow += -pw;
      }
    }
    for (int k = 0; k < A.ckt.nel; k++) {  // model-supplied (negative resistors, controlled sources)
      const int32_t* e = el(k);
      if ((e[ET_KIND] == K_R && L.P(e, 0) < 0) || is_controlled(e[ET_KIND])) {
        OpPt o = operating(L, e, x_);
        double ms = -o.power;
        if (ms != 0) { sup += ms; scale += fabs(ms); }
      }
    }
    for (int q = 0; q < A.ckt.n_rx; q++) {  // storage of reactive devices
      const int32_t* e = et_ + (int64_t)A.ckt.rx[q] * ET_W;
      OpPt o = operating(L, e, x_);
      stor += o.storage;
      scale += fabs(o.storage);
    }
    int err = ST_OK;
    if (scale > cd[CD_PEAK]) cd[CD_PEAK] = scale;
    double res = fabs(inflow + sup - outflow - diss - stor);
    if (!imposed && res > fmax(1e-21, self().roundoff())) {
      if (res > cd[CD_WORST]) cd[CD_WORST] = res;
      if (res / cd[CD_PEAK] > 1e-3) err = ST_LEDGER;
    }
    (void)t;
    return err;
  }

  // ---- solve request (Sim.solve, analysis/op.go:289-322): stamp now, factor
  // in the next LU launch.
  KH_HD void request_solve() {
    const uint64_t t0 = KH_CLOCK();
    self().vcopy(A.xold, A.x);
    self().zero_system();
    self().load_all();
    self().diag_terms();
    ci[CI_NEED_SOLVE] = 1;
    if (L.grew && !A.ckt.dense) ci[CI_WAIT] = WAIT_PATTERN;  // the dense LU has no pattern
    cd[CD_CYC_STAMP] += (double)(KH_CLOCK() - t0);
  }
  // Sim.solve's terms after the devices' (analysis/op.go:300-318): the
  // pseudo-transient diagonal +-1/h with its history, the gmin-stepping
  // diagonal.
  KH_HD void diag_terms() {
    double ptG = cd[CD_PTG];
    if (ptG > 0) {
      for (int i = 0; i < n; i++) {
        int s = diag_slot(i);
        if (A.ckt.isV[i]) { add_diag(s, ptG); X(A.rhs, i) += ptG * X(A.ptprev, i); }
        else { add_diag(s, -ptG); X(A.rhs, i) -= ptG * X(A.ptprev, i); }
      }
    }
    double g = cd[CD_DIAG_GMIN];
    if (g > 0) {
      for (int i = 0; i < n; i++)
        if (A.ckt.isV[i]) add_diag(diag_slot(i), g);
    }
  }
  KH_HD void zero_system() {
    for (int i = 0; i < n; i++) X(A.rhs, i) = 0;
    for (int s = 0; s < A.ckt.nslots; s++) A.val[s] = 0;
  }
  // Vector helpers over the n unknowns.
  KH_HD void vzero(double* v) { for (int i = 0; i < n; i++) v[i] = 0; }
  KH_HD void vcopy(double* d, const double* s) { copy_n(d, s, n); }
  // pseudoTransient's steady-state test (analysis/newton.go:291-300).
  KH_HD bool pt_moved() {
    for (int i = 0; i < n; i++) {
      double tol = A.ckt.isV[i] ? P.vnTol : P.absTol;
      double xv = X(A.x, i);
      if (fabs(xv - X(A.ptprev, i)) > P.relTol * fabs(xv) + tol) return true;
    }
    return false;
  }
  // polish's stopping test (analysis/op.go:262-267).
  KH_HD bool polish_done() {
    bool done = true;
    for (int i = 0; i < n; i++) {
      double xv = X(A.x, i);
      if (fabs(xv - X(A.xold, i)) > 1e-12 * fmax(1.0, fabs(xv))) done = false;
    }
    return done;
  }
  // SetInitialCondition of every C and L (analysis/tran.go:100-108).
  KH_HD void uic_init() {
    for (int k = 0; k < A.ckt.nel; k++) {
      const int32_t* e = el(k);
      if (e[ET_KIND] == K_C) { L.S(e, 3) = L.P(e, 2) != 0 ? L.P(e, 1) : 0.0; L.S(e, 4) = 0; }
      if (e[ET_KIND] == K_L) { L.S(e, 0) = 0; L.S(e, 1) = L.P(e, 2) != 0 ? L.P(e, 1) : 0.0; }
    }
  }
  KH_HD void sweep_store(int point) {
    for (int i = 0; i < n; i++) A.sw_x[(int64_t)point * n + i] = X(A.x, i);
    if (A.sw_i) {  // DCSweepResult.I: each element's Operating current after Accept (analysis/dc.go:121-123)
      const Lane L = this->L;
      const int nel_ = A.ckt.nel;
      for (int k = 0; k < nel_; k++) A.sw_i[(int64_t)point * nel_ + k] = operating(L, el(k), A.x).current;
    }
  }
  // Slot of the diagonal entry (i, i): the device slot when a device stamps
  // it, else a slot reserved after the device slots.
  KH_HD int diag_slot(int i) const { return A.ckt.diag_slot[i]; }
  KH_HD void add_diag(int s, double v) {
    int64_t ix = s;
    A.val[ix] = A.val[ix] + v;
    if (!A.present[ix]) {
      A.present[ix] = 1;
      L.grew = true;
    }
  }

  KH_HD void start_newton(int maxIter, int caller) {
    ci[CI_NT_IT] = 0;
    ci[CI_NT_MAX] = maxIter;
    ci[CI_CALLER] = caller;
    cd[CD_GMIN] = P.gmin;  // s.lc.Gmin = o.Gmin, newton.go:37
    request_solve();
  }

  // After a factor+solve: the body of the Newton loop (analysis/newton.go:38-56).
  KH_HD void after_solve(int e) {
    ci[CI_ITERS]++;
    ci[CI_SOLVES]++;
    if (ci[CI_CALLER] == CALLER_POLISH) { polish_step(e); return; }
    if (e != ST_OK) { newton_done(e); return; }
    if (!A.ckt.nonlin) { newton_done(ST_OK); return; }
    ci[CI_NT_IT]++;
    if (ci[CI_INITJCT]) {
      ci[CI_INITJCT] = 0;
      if (ci[CI_NT_IT] >= ci[CI_NT_MAX]) newton_done(ST_NOCONV);
      else request_solve();
      return;
    }
    const uint64_t t0 = KH_CLOCK();
    const bool conv = self().converged();
    cd[CD_CYC_CONV] += (double)(KH_CLOCK() - t0);
    if (conv) { newton_done(ST_OK); return; }
    if (ci[CI_NT_IT] >= ci[CI_NT_MAX]) newton_done(ST_NOCONV);
    else request_solve();
  }

  KH_HD void newton_done(int r) {
    switch (ci[CI_CALLER]) {
      case CALLER_DC_PLAIN: dc_plain_done(r); break;
      case CALLER_GMIN: gmin_done(r); break;
      case CALLER_SRC0: src0_done(r); break;
      case CALLER_SRC: src_done(r); break;
      case CALLER_PT: pt_done(r); break;
      case CALLER_PT_FINAL: pt_end(r); break;
      case CALLER_UIC: uic_done(r); break;
      case CALLER_TRAN: tran_newton_done(r); break;
    }
  }

  // ---- dcSolve, analysis/newton.go:95-128 -------------------------------------
  KH_HD void dc_begin(int ret) {
    ci[CI_DC_RET] = ret;
    ci[CI_INITJCT] = ci[CI_WARM] ? 0 : 1;
    if (!ci[CI_WARM] && A.nodeset) {
      self().apply_nodeset();
      ci[CI_INITJCT] = 0;
    }
    start_newton(P.itl1, CALLER_DC_PLAIN);
  }
  // applyNodeSet, analysis/newton.go:221-243: x <- the guess (node k is
  // unknown k-1, the rest 0), then every nonlinear device seeds its limiting
  // state from it.
  KH_HD void apply_nodeset() {
    self().vzero(A.x);
    for (int i = 0; i < A.ckt.numNodes - 1; i++) X(A.x, i) = A.nodeset[i];
    for (int q = 0; q < A.ckt.n_nl; q++) {
      const int32_t* e = el(A.ckt.nl[q]);
      switch (e[ET_KIND]) {
        case K_D: if (KH_HAS(K_D)) accept_D(L, e, A.x, 0, MODE_OP); break;
        case K_M: if (KH_HAS(K_M)) accept_M(L, e, A.x, 0, MODE_OP); break;
        case K_Q: if (KH_HAS(K_Q)) accept_Q(L, e, A.x, 0, MODE_OP); break;
      }
    }
  }
  KH_HD void dc_return(int r) {
    if (ci[CI_DC_RET] == DCRET_OP) op_after_dc(r);
    else tran_after_dc(r);
  }
  KH_HD void dc_plain_done(int r) {
    if (r == ST_OK || !A.ckt.nonlin || !retryable(r)) {
      if (r == ST_OK) ci[CI_WARM] = 1;
      dc_return(r);
      return;
    }
    ci[CI_FIRST_ERR] = r;
    dc_try_gmin();
  }
  KH_HD void dc_try_gmin() {
    if (!P.noGmin) gmin_begin();
    else dc_try_source();
  }
  KH_HD void dc_try_source() {
    if (!P.noSource) src_begin();
    else dc_try_pt();
  }
  KH_HD void dc_try_pt() {
    if (!P.noPT) pt_begin();
    else dc_return(ci[CI_FIRST_ERR]);
  }

  // gminStepping, analysis/newton.go:134-176
  KH_HD void gmin_begin() {
    self().vzero(A.x);
    self().vzero(A.xsave);
    ci[CI_INITJCT] = 1;
    cd[CD_GM_G] = 1e-2;
    cd[CD_GM_FACTOR] = 10.0;
    ci[CI_GM_OK] = 0;
    ci[CI_GM_K] = 0;
    cd[CD_DIAG_GMIN] = cd[CD_GM_G];
    start_newton(P.itl1, CALLER_GMIN);
  }
  KH_HD void gmin_done(int r) {
    double g = cd[CD_GM_G], factor = cd[CD_GM_FACTOR];
    if (r == ST_OK) {
      self().vcopy(A.xsave, A.x);
      if (g == 0) { gmin_end(ST_OK); return; }
      ci[CI_GM_OK] = 1;
      factor = fmin(factor * sqrt(factor), 10.0);
      if (g / factor < P.gmin) g = 0;
      else g /= factor;
      cd[CD_GM_G] = g; cd[CD_GM_FACTOR] = factor;
      if (++ci[CI_GM_K] >= 200) { gmin_end(ST_NOCONV); return; }
      cd[CD_DIAG_GMIN] = g;
      start_newton(P.itl1, CALLER_GMIN);
      return;
    }
    if (!retryable(r)) { gmin_end(r); return; }
    if (!ci[CI_GM_OK]) { gmin_end(ST_NOCONV); return; }
    self().vcopy(A.x, A.xsave);
    double prev = g * factor;
    if (g == 0) prev = P.gmin;
    factor = sqrt(factor);
    if (factor < 1.00005) { gmin_end(ST_NOCONV); return; }
    g = prev / factor;
    cd[CD_GM_G] = g; cd[CD_GM_FACTOR] = factor;
    if (++ci[CI_GM_K] >= 200) { gmin_end(ST_NOCONV); return; }
    cd[CD_DIAG_GMIN] = g;
    start_newton(P.itl1, CALLER_GMIN);
  }
  KH_HD void gmin_end(int r) {
    cd[CD_DIAG_GMIN] = 0;
    if (r == ST_OK) { ci[CI_WARM] = 1; dc_return(ST_OK); return; }
    dc_try_source();
  }

  // sourceStepping, analysis/newton.go:180-214
  KH_HD void src_begin() {
    self().vzero(A.x);
    self().vzero(A.xsave);
    ci[CI_SRCSTEP] = 1;
    cd[CD_SS_DONE] = 0;
    cd[CD_SS_STEP] = 0.1;
    cd[CD_SRCFACT] = 0;
    start_newton(P.itl1, CALLER_SRC0);
  }
  KH_HD void src0_done(int r) {
    if (r != ST_OK) { src_end(r); return; }
    self().vcopy(A.xsave, A.x);
    ci[CI_SS_K] = 0;
    src_next();
  }
  KH_HD void src_next() {
    double f = fmin(cd[CD_SS_DONE] + cd[CD_SS_STEP], 1.0);
    cd[CD_SS_F] = f;
    cd[CD_SRCFACT] = f;
    start_newton(P.itl1, CALLER_SRC);
  }
  KH_HD void src_done(int r) {
    double f = cd[CD_SS_F];
    if (r == ST_OK) {
      if (f == 1) { src_end(ST_OK); return; }
      self().vcopy(A.xsave, A.x);
      cd[CD_SS_DONE] = f;
      cd[CD_SS_STEP] = fmin(cd[CD_SS_STEP] * 1.5, 0.5);
      if (++ci[CI_SS_K] >= 1000) { src_end(ST_NOCONV); return; }
      src_next();
      return;
    }
    if (!retryable(r)) { src_end(r); return; }
    self().vcopy(A.x, A.xsave);
    cd[CD_SS_STEP] /= 2;
    if (cd[CD_SS_STEP] < 1e-3) { src_end(ST_NOCONV); return; }
    if (++ci[CI_SS_K] >= 1000) { src_end(ST_NOCONV); return; }
    src_next();
  }
  KH_HD void src_end(int r) {
    ci[CI_SRCSTEP] = 0;
    cd[CD_SRCFACT] = 0;
    if (r == ST_OK) { ci[CI_WARM] = 1; dc_return(ST_OK); return; }
    dc_try_pt();
  }

  // pseudoTransient, analysis/newton.go:260-311
  KH_HD void pt_begin() {
    self().vzero(A.x);
    self().vzero(A.ptprev);
    ci[CI_INITJCT] = 1;
    cd[CD_PT_H] = 1e-6;
    ci[CI_PT_QUIET] = 0;
    ci[CI_PT_K] = 0;
    cd[CD_PTG] = 1 / cd[CD_PT_H];
    start_newton(P.itl1, CALLER_PT);
  }
  KH_HD void pt_done(int r) {
    if (r != ST_OK) {
      if (!retryable(r)) { pt_end(r); return; }
      self().vcopy(A.x, A.ptprev);
      cd[CD_PT_H] /= 4;
      if (cd[CD_PT_H] < 1e-15) { pt_end(ST_NOCONV); return; }
      if (++ci[CI_PT_K] >= 2000) { pt_end(ST_NOCONV); return; }
      cd[CD_PTG] = 1 / cd[CD_PT_H];
      start_newton(P.itl1, CALLER_PT);
      return;
    }
    const bool moved = self().pt_moved();
    self().vcopy(A.ptprev, A.x);
    if (!moved) ci[CI_PT_QUIET]++;
    else ci[CI_PT_QUIET] = 0;
    if (ci[CI_PT_QUIET] >= 3 && cd[CD_PT_H] >= 1e6) {
      cd[CD_PTG] = 0;
      start_newton(P.itl1, CALLER_PT_FINAL);
      return;
    }
    cd[CD_PT_H] = fmin(cd[CD_PT_H] * 2, 1e12);
    if (++ci[CI_PT_K] >= 2000) { pt_end(ST_NOCONV); return; }
    cd[CD_PTG] = 1 / cd[CD_PT_H];
    start_newton(P.itl1, CALLER_PT);
  }
  KH_HD void pt_end(int r) {
    cd[CD_PTG] = 0;
    if (r == ST_OK) { ci[CI_WARM] = 1; dc_return(ST_OK); return; }
    dc_return(ci[CI_FIRST_ERR]);
  }

  // ---- OP, analysis/op.go:340-357 (+ polish, :251-271) -------------------------
  KH_HD void op_begin() {
    ci[CI_MODE] = MODE_OP;
    cd[CD_TIME] = 0;
    dc_begin(DCRET_OP);
  }
  KH_HD void op_after_dc(int r) {
    if (P.analysis == AN_DC) { sweep_point_done(r); return; }
    if (r != ST_OK) { finish(r); return; }
    self().accept(0, MODE_OP);
    double rr = self().ledger_op();
    cd[CD_LEDGER] = rr;
    if (!P.skipEnergy && rr > P.ledgerTol) { finish(ST_LEDGER); return; }
    if (P.analysis == AN_OP_POLISH && A.ckt.nonlin) {
      ci[CI_PL_K] = 0;
      ci[CI_CALLER] = CALLER_POLISH;
      request_solve();
      return;
    }
    finish(ST_OK);
  }
  KH_HD void polish_step(int e) {
    if (e != ST_OK) { finish(e); return; }
    const bool done = self().polish_done();
    ci[CI_PL_K]++;
    if (done || ci[CI_PL_K] >= 2) {
      self().accept(0, MODE_OP);
      finish(ST_OK);
      return;
    }
    request_solve();
  }

  // ---- DC sweep, analysis/dc.go:51-130 ------------------------------------------
  // Every point overrides the swept sources' DC values (SetDCOverride,
  // device/source.go:179-203: in ModeOP the source takes the override, which is
  // the parameter pair dc = v, hasDC = 1) and runs Sim.OP on the same circuit
  // state, so every point after the first Newton-starts warm from the last.
  KH_HD void sweep_set_sources() {
    const int idx[2] = {ci[CI_SW_I], ci[CI_SW_J]};
    for (int s = 0; s < P.swN; s++) {
      const int32_t* e = el(P.swElem[s]);
      double* par = const_cast<double*>(A.par);
      par[e[ET_POFF] + 0] = P.swStart[s] + (double)idx[s] * P.swStep[s];
      par[e[ET_POFF] + 1] = 1.0;
    }
  }
  KH_HD void sweep_begin() {
    ci[CI_SW_I] = 0;
    ci[CI_SW_J] = 0;
    ci[CI_SW_DONE] = 0;
    cd[CD_MAXREL] = 0;
    sweep_point_begin();
  }
  KH_HD void sweep_point_begin() {
    sweep_set_sources();
    ci[CI_SW_IT0] = ci[CI_ITERS];
    ci[CI_MODE] = MODE_OP;
    cd[CD_TIME] = 0;
    dc_begin(DCRET_OP);
  }
  KH_HD void sweep_point_done(int r) {
    const int point = ci[CI_SW_J] * P.swCount[0] + ci[CI_SW_I];
    A.sw_it[point] = ci[CI_ITERS] - ci[CI_SW_IT0];
    if (r != ST_OK) { finish(r); return; }
    self().accept(0, MODE_OP);
    double rr = self().ledger_op();
    cd[CD_LEDGER] = rr;
    if (rr > cd[CD_MAXREL]) cd[CD_MAXREL] = rr;
    if (!P.skipEnergy && rr > P.ledgerTol) { finish(ST_LEDGER); return; }
    self().sweep_store(point);
    ci[CI_SW_DONE]++;
    if (++ci[CI_SW_I] == P.swCount[0]) {
      ci[CI_SW_I] = 0;
      if (P.swN < 2 || ++ci[CI_SW_J] == P.swCount[1]) { finish(ST_OK); return; }
    }
    sweep_point_begin();
  }

  // ---- transient, analysis/tran.go:95-231 --------------------------------------
  KH_HD void tran_begin() {
    if (P.uic) {
      self().uic_init();
      ci[CI_MODE] = MODE_UICINIT;
      cd[CD_TIME] = 0;
      ci[CI_INITJCT] = 1;
      start_newton(P.itl1, CALLER_UIC);
      return;
    }
    ci[CI_MODE] = MODE_TRANOP;
    cd[CD_TIME] = 0;
    dc_begin(DCRET_TRAN);
  }
  KH_HD void uic_done(int r) {
    if (r != ST_OK) { finish(r); return; }
    self().accept(0, MODE_UICINIT);
    self().record(0, true);
    tran_init_loop();
  }
  KH_HD void tran_after_dc(int r) {
    if (r != ST_OK) { finish(r); return; }
    self().accept(0, MODE_TRANOP);
    int le = self().record(0, false);
    if (le != ST_OK && !P.skipEnergy) { finish(le); return; }
    tran_init_loop();
  }
  KH_HD void tran_init_loop() {
    cd[CD_TR_T] = 0;
    cd[CD_TR_H] = fmin(P.maxStep, fmin(P.step, P.stop / 100)) / 10;
    ci[CI_TR_ORDER] = BE;
    self().vcopy(A.xacc, A.x);
    cd[CD_MINSTEP_SEEN] = kInf;
    cd[CD_MAXSTEP_SEEN] = 0;
    tran_loop_top();
  }
  KH_HD void tran_loop_top() {
    double t = cd[CD_TR_T], h = cd[CD_TR_H];
    double eps = 1e-12 * P.stop;
    if (!(t < P.stop - eps)) { tran_finish(); return; }
    if (ci[CI_ACCEPTED] >= P.maxSteps) { finish(ST_MAXSTEPS); return; }
    double nb = self().next_breakpoint(t);
    double tn = t + h;
    int hit = 0;
    if (tn >= nb - eps) { tn = nb; hit = 1; }
    if (tn >= P.stop - eps) { tn = P.stop; hit = 0; }
    h = tn - t;
    cd[CD_TR_H] = h;
    cd[CD_TR_TN] = tn;
    ci[CI_TR_HIT] = hit;
    ci[CI_MODE] = MODE_TRAN;
    cd[CD_TIME] = tn;
    cd[CD_H] = h;
    ci[CI_METHOD] = ci[CI_TR_ORDER];
    start_newton(P.itl4, CALLER_TRAN);
  }
  KH_HD bool reject(double hNew) {  // tran.go:136-143; true if the step became too small
    ci[CI_REJECTED]++;
    self().vcopy(A.x, A.xacc);
    return hNew < P.minStep;
  }
  KH_HD void tran_newton_done(int r) {
    double h = cd[CD_TR_H], tn = cd[CD_TR_TN];
    int order = ci[CI_TR_ORDER];
    if (r != ST_OK) {
      if (!retryable(r) || !A.ckt.nonlin) { finish(r); return; }
      h /= 8;
      cd[CD_TR_H] = h;
      if (reject(h)) { finish(ST_TSTEP); return; }
      ci[CI_TR_ORDER] = BE;
      tran_loop_top();
      return;
    }
    uint64_t t0 = KH_CLOCK();
    double proposal = self().truncate(tn, h, order);
    cd[CD_CYC_LTE] += (double)(KH_CLOCK() - t0);
    if (proposal < 0.9 * h) {
      h = fmax(proposal, h / 8);
      cd[CD_TR_H] = h;
      if (reject(h)) { finish(ST_TSTEP); return; }
      tran_loop_top();
      return;
    }
    t0 = KH_CLOCK();
    self().accept(tn, MODE_TRAN);
    const uint64_t t1 = KH_CLOCK();
    int le = self().record(tn, false);
    cd[CD_CYC_ACCEPT] += (double)(t1 - t0);
    cd[CD_CYC_LEDGER] += (double)(KH_CLOCK() - t1);
    if (le != ST_OK && !P.skipEnergy) { finish(le); return; }
    self().vcopy(A.xacc, A.x);
    ci[CI_ACCEPTED]++;
    cd[CD_MINSTEP_SEEN] = fmin(cd[CD_MINSTEP_SEEN], h);
    cd[CD_MAXSTEP_SEEN] = fmax(cd[CD_MAXSTEP_SEEN], h);
    cd[CD_TR_T] = tn;
    double next = fmin(proposal, 2 * h);
    next = fmin(next, P.maxStep);
    ci[CI_TR_ORDER] = P.method;
    if (ci[CI_TR_HIT]) {
      ci[CI_TR_ORDER] = BE;
      double nb2 = self().next_breakpoint(tn);
      next = fmin(next, 0.1 * fmin(P.maxStep, nb2 - tn));
    }
    cd[CD_TR_H] = fmax(next, P.minStep);
    tran_loop_top();
  }
  KH_HD void tran_finish() {
    if (cd[CD_PEAK] > 0) {
      cd[CD_MAXREL] = cd[CD_WORST] / cd[CD_PEAK];
      if (cd[CD_MAXREL] > 1e-9 && !P.skipEnergy) { finish(ST_LEDGER); return; }
    }
    finish(ST_OK);
  }

  // ---- one control step of the lane ------------------------------------------
  KH_HD void step() {
    if (ci[CI_PHASE] == PH_DONE) return;
    if (ci[CI_WAIT]) return;  // waiting for the host (a grown pattern, more L/U capacity)
    if (ci[CI_PHASE] == PH_INIT) {
      ci[CI_PHASE] = PH_RUN;
      if (P.analysis == AN_TRAN) tran_begin();
      else if (P.analysis == AN_DC) sweep_begin();
      else op_begin();
      return;
    }
    if (!ci[CI_NEED_SOLVE]) return;
    int e = A.lu_status[0];
    ci[CI_NEED_SOLVE] = 0;
    if (e == ST_CAPACITY) { finish(ST_CAPACITY); return; }
    after_solve(e);
  }
};

// The state machine on one lane: every loop runs serially.
struct Ctl : CtlT<Ctl> {
  KH_HD Ctl(const CtlArgs& a) : CtlT<Ctl>(a) {}
};

}  // inline namespace kinds_<KH_KINDS>
}  // namespace kh
