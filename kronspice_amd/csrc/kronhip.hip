// kronhip.hip — gfx950 kernels and the C ABI of include/kronhip.h.
//
// Kernels:
//   k_persist — each circuit alternates factor+solve (kh_lu.h) with its
//               analysis state machine (kh_control.h) until the analysis ends
//               or its stamps need a new ordering; one launch covers every
//               topology of an uneven batch (per-circuit storage, kh_group.h)
//   k_lu      — factor + solve only, one lane per system with the system index
//               fastest (SoA): the per-system seam and the batched LU
//   k_hash    — fingerprints the sparsity pattern of lanes whose stamps
//               created new entries, for the host's symbolic phase
//   k_tear    — torn solves (kh_tear.h), one system per lane; k_tear_untorn
//               solves the systems they refuse with the untorn LU
// The host relaunches k_persist until every circuit is done and runs the
// symbolic phase (kh_symbolic.cpp) once per distinct pattern in between.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <map>
#include <thread>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/kronhip.h"
#define KH_CLOCK() ((uint64_t)clock64())
#include "kh_ac.h"
#include "kh_group.h"
#include "kh_persist.h"
#include "kh_report.h"
#include "kh_sched.h"
#include "kh_symbolic.h"
#include "kh_tear.h"
#include "kh_wave.h"

using namespace kh;

namespace {

thread_local std::string g_last_error;

kh_status fail(kh_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

#define KH_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(KH_HIP_ERROR, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)


// ---- kernels ----------------------------------------------------------------------
// k_persist lives in kh_persist.hip, compiled once per register budget
// (kh_persist.h); the host launches it through kh_launch_persist.

// Factor + solve the lanes that requested a solve and have a structure.
// need == nullptr: every lane.
__global__ void __launch_bounds__(kBlock) k_lu(LuArgs a, const int32_t* need, const int32_t* wait, int32_t* status) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.B) return;
  if (need && (!need[c] || wait[c])) return;
  int st = lu_factor(a, c);
  if (st == LU_OK) st = lu_solve(a, c);
  status[c] = st;
}

// Mixed precision factor + solve of every lane (kh_lu_batch_solve_mixed):
// FP32 factors and FP64 refinement; a lane the mixed path cannot certify is
// redone in FP64 exactly as k_lu does (passes[c] = 0 then).
__global__ void __launch_bounds__(kBlock) k_lu_mixed(LuArgs a, int32_t* status, int32_t* passes) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.B) return;
  int it = 0;
  int st = lu_factor_mixed(a, c);
  if (st == LU_OK) st = lu_solve_mixed(a, c, &it);
  if (st == LU_UNSURE) {
    it = 0;
    st = lu_factor(a, c);
    if (st == LU_OK) st = lu_solve(a, c);
  }
  status[c] = st;
  if (passes) passes[c] = it;
}

// FNV-1a over the present slot ids of each waiting circuit (contiguous
// per-circuit storage).
__global__ void __launch_bounds__(kBlock) k_hash(const uint8_t* present, int32_t nslots, int64_t B,
                                                 const int32_t* ci, uint64_t* hash) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B) return;
  if (!ci[c * NCI + CI_WAIT]) return;
  const uint8_t* pr = present + c * nslots;
  uint64_t h = 1469598103934665603ull;
  for (int s = 0; s < nslots; s++) {
    if (pr[s]) {
      h ^= (uint64_t)(s + 1);
      h *= 1099511628211ull;
    }
  }
  hash[c] = h;
}

// dst[c*rows + r] = src[r*cols + c]: SoA (caller layout) <-> per-circuit layout.
template <typename T>
__global__ void __launch_bounds__(256) k_transpose(const T* src, T* dst, int64_t rows, int64_t cols) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= rows * cols) return;
  int64_t r = k / cols, c = k % cols;
  dst[c * rows + r] = src[k];
}

// Per-circuit results into the caller's SoA arrays.
__global__ void __launch_bounds__(256) k_outputs(const int32_t* ci, const double* cd, int64_t B, int32_t* status,
                                                 int32_t* iters, int32_t* acc, int32_t* rej, double* ledger,
                                                 int32_t* solves, int32_t* points) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B) return;
  const int32_t* w = ci + c * NCI;
  if (status) status[c] = w[CI_STATUS];
  if (iters) iters[c] = w[CI_ITERS];
  if (acc) acc[c] = w[CI_ACCEPTED];
  if (rej) rej[c] = w[CI_REJECTED];
  if (ledger) ledger[c] = cd[c * NCD + CD_MAXREL];
  if (solves) solves[c] = w[CI_SOLVES];
  if (points) points[c] = w[CI_SW_DONE];
}

// ---- element reports (kh_report.h) ----------------------------------------------------
static_assert(KH_REPORT_FIELDS == RP_FIELDS && KH_RP_CHANNEL == RP_CHANNEL && KH_RP_ENERGY == RP_ENERGY,
              "report fields of kronhip.h and kh_report.h");
static_assert(KH_LEDGER_FIELDS == LG_FIELDS && KH_LG_RELATIVE == LG_RELATIVE && KH_LG_ROUNDOFF == LG_ROUNDOFF,
              "ledger fields of kronhip.h and kh_report.h");
// One element of one circuit per thread: its output.Result entry at the
// circuit's final solution and device state, rep[c][k][RP_FIELDS].
__global__ void __launch_bounds__(256) k_report(const Group g, double* rep) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nel = g.ctl.ckt.nel;
  if (t >= g.ctl.B * nel) return;
  const int64_t c = t / nel;
  const int k = (int)(t % nel);
  CtlArgs a;
  LuArgs l;
  circuit_view(g, c, a, l);
  Lane L = {};
  L.n = a.ckt.n;
  L.et = a.ckt.et;
  L.par = a.par;
  L.st = a.st;
  report_element(L, a.ckt.et + (int64_t)k * ET_W, a.x, rep + t * RP_FIELDS);
}

// One circuit per thread: the ledger (em.BuildLedger) over its element
// reports, with the round-off floor of its last stamped system,
// led[c][LG_FIELDS].
__global__ void __launch_bounds__(64) k_report_ledger(const Group g, const double* rep, double* led) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ctl.B) return;
  CtlArgs a;
  LuArgs l;
  circuit_view(g, c, a, l);
  Ctl ctl(a);
  const double ro = ctl.roundoff();
  const int nel = a.ckt.nel;
  fold_ledger(a.ckt.et, nel, rep + c * nel * RP_FIELDS, ro, led + c * LG_FIELDS);
}

// ---- AC systems (kh_ac.h) ----------------------------------------------------------
// Lane l of a chunk of AC systems is frequency f0 + l % nfc of circuit
// l / nfc: consecutive lanes are one circuit's consecutive frequencies (whose
// systems tend to share pivot rows: the compiled schedules, kh_sched.h, run a
// workgroup of them in lockstep).
struct AcArgs {
  Circ ckt;
  const double* par;     // [B][npar] per-circuit parameters
  const double* st;      // [B][nstate] device state accepted at the operating point
  const int32_t* ci;     // [B][NCI] control words (status of the operating point)
  int32_t npar, nstate;
  int64_t B, BL;         // circuits; lanes of this chunk
  int32_t f0;            // first frequency of the chunk
  int32_t nfc;           // frequencies in the chunk (BL = B * nfc)
  const double* freqs;   // [nf]
  kcplx* val;            // [nslots][BL]
  kcplx* rhs;            // [n][BL]
  uint8_t* present;      // [B][nslots] stamped-slot marks (first chunk), or null
  int32_t* lstatus;      // [BL]: -1 skipped (no operating point), else status
  double* lrr;           // [BL] ledger relative residual
  int32_t nf;            // frequencies of the sweep
  double* iout;          // optional [nf][nel][B] complex element currents
  double* lout;          // optional [nf][AL_FIELDS][B] per-frequency ledgers
};

KH_HD inline ACLane ac_lane(const AcArgs& a, int64_t l) {
  const int64_t c = l / a.nfc;
  const int32_t f = a.f0 + (int32_t)(l % a.nfc);
  ACLane A;
  A.dev = Lane{};
  A.dev.n = a.ckt.n;
  A.dev.et = a.ckt.et;
  A.dev.par = a.par + c * a.npar;
  A.dev.st = const_cast<double*>(a.st) + c * a.nstate;
  A.l = l;
  A.BL = a.BL;
  A.val = a.val;
  A.rhs = a.rhs;
  A.present = nullptr;
  A.omega = 2 * 3.14159265358979323846 * a.freqs[f];
  return A;
}

// Stamps every lane's AC system (LoadAC of each element at omega).
__global__ void __launch_bounds__(kBlock) k_ac_stamp(AcArgs a) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= a.BL) return;
  for (int s = 0; s < a.ckt.nslots; s++) a.val[(int64_t)s * a.BL + l] = kcplx{0, 0};
  for (int i = 0; i < a.ckt.n; i++) a.rhs[(int64_t)i * a.BL + l] = kcplx{0, 0};
  const int64_t c = l / a.nfc;
  if (a.ci[c * NCI + CI_STATUS] != ST_OK) {
    a.lstatus[l] = -1;
    return;
  }
  ACLane A = ac_lane(a, l);
  if (a.present && l < a.B) A.present = a.present + c * a.ckt.nslots;
  ac_load(A, a.ckt.nel);
  a.lstatus[l] = ST_OK;
}

// Factors and solves every stamped lane (complex GP LU), evaluates its
// ledger (analysis/ac.go:148-190) and writes its solution to the caller's
// [nf][n][B] complex array.
// precision 1: complex64 factors with complex128 refinement (kh_lu.h), the
// complex128 path for every system it cannot certify; mixl[l] receives the
// lane's complex64 solves, -1 where it was redone (summed on the host).
// lock (or null): the lanes' solutions are already in lu.x — 1 where the
// compiled schedules solved the lane, else lustat[l] is the per-lane LU's
// status.
__global__ void __launch_bounds__(kBlock) k_ac_solve(AcArgs a, LuArgsT<kcplx> lu, double ledger_tol,
                                                     int32_t skip_energy, double* xout, int32_t precision,
                                                     int32_t* mixl, const int32_t* lock, const int32_t* lustat) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= a.BL || a.lstatus[l] < 0) return;
  int st;
  if (lock) {
    st = lock[l] ? (int)LU_OK : lustat[l];
  } else if (a.ckt.dense) {
    st = lu_dense(lu, l, a.ckt.slot_i, a.ckt.slot_j);
  } else {
    st = LU_UNSURE;
    if (precision) {
      int passes = 0;
      st = lu_factor_mixed(lu, l);
      if (st == LU_OK) st = lu_solve_mixed(lu, l, &passes);
      mixl[l] = st == LU_OK ? passes : (st == LU_UNSURE ? -1 : 0);
    }
    if (st == LU_UNSURE) {
      st = lu_factor(lu, l);
      if (st == LU_OK) st = lu_solve(lu, l);
    }
  }
  if (st != LU_OK) {
    a.lstatus[l] = st;
    return;
  }
  const int n = a.ckt.n;
  const int64_t BL = a.BL;
  ACLane A = ac_lane(a, l);
  kcplx sumS;
  double scaleS;
  const int64_t cc = l / a.nfc, ff = a.f0 + l % a.nfc;
  kcplx* icur = a.iout ? (kcplx*)a.iout + (ff * a.ckt.nel) * a.B + cc : nullptr;
  double led[AL_FIELDS];
  ac_power(A, a.ckt.nel, lu.x, &sumS, &scaleS, icur, a.B, a.lout ? led : nullptr);
  if (a.lout)
    for (int q = 0; q < AL_FIELDS; q++) a.lout[(ff * AL_FIELDS + q) * a.B + cc] = led[q];
  // Round-off floor (output.Ledger.Roundoff): 64 eps sum |x_i| (sum_j |A_ij||x_j| + |b_i|).
  // AbsProduct row by row (ckt.ro_ptr / ro_slot: a row's slots in slot order
  // for CSparse, solver/sparse.go:519-527, in column order for CDense,
  // solver/cdense.go:57-67), each row's sum kept in a register: the same
  // additions in the same order as accumulating over all slots.
  double ro = 0;
  for (int i = 0; i < n; i++) {
    double acc = 0;
    for (int p = a.ckt.ro_ptr[i]; p < a.ckt.ro_ptr[i + 1]; p++) {
      const int s = a.ckt.ro_slot[p];
      acc += lu_mag(a.val[(int64_t)s * BL + l]) * lu_mag(lu.x[(int64_t)a.ckt.slot_j[s] * BL + l]);
    }
    ro += lu_mag(lu.x[(int64_t)i * BL + l]) * (acc + lu_mag(a.rhs[(int64_t)i * BL + l]));
  }
  ro *= 64 * 0x1p-52;
  const double mag = go_hypot(sumS.re, sumS.im);
  const double rr = (scaleS > 0 && mag > ro) ? mag / scaleS : 0.0;
  a.lrr[l] = rr;
  a.lstatus[l] = (!skip_energy && rr > ledger_tol) ? ST_LEDGER : ST_OK;
  const int64_t c = l / a.nfc, f = a.f0 + l % a.nfc;
  for (int i = 0; i < n; i++) {
    const kcplx v = lu.x[(int64_t)i * BL + l];
    const int64_t o = ((f * n + i) * a.B + c) * 2;
    xout[o] = v.re;
    xout[o + 1] = v.im;
  }
}

// Lockstep factor + solve along compiled schedules (kh_sched.h): workgroup w
// runs its lanes along views[wg_view[w]] (-1: none this round); lock[l] = 1
// where a lane succeeded, else dep[l] is the column at which its pivot rule
// departed from the schedule (-2: a check the per-lane LU must report). A lane
// already locked, without an operating point or of another pattern is skipped.
// compact: 1, the working columns are compact (sched_lu_lds) in LDS; 2,
// compact in device memory, workgroup w's block at xg + w * xg_block doubles
// (small enough to stay in cache); 0, dense in device memory (sched_lu).
__global__ void __launch_bounds__(kBlock) k_ac_sched(AcArgs a, LuArgsT<kcplx> lu, const SchedView* views,
                                                     const int32_t* wg_view, int32_t* lock, int32_t* dep,
                                                     int32_t* dpiv, int32_t compact, double* xg, int64_t xg_block) {
  extern __shared__ __align__(16) char lds[];
  const int32_t v = __builtin_amdgcn_readfirstlane(wg_view[blockIdx.x]);  // uniform: one per workgroup
  if (v < 0) return;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= a.BL || a.lstatus[l] < 0 || lock[l]) return;
  const SchedView s = ((SchedPtr<SchedView>)views)[v];
  if (lu.pid[l] != s.pid) return;
  int d = -2, dp = -1, st;
  const SchedLane<kcplx> P = sched_lane_blocked(lu, l, (int)blockDim.x);
  if (compact == 1) {
    using LdsD = __attribute__((address_space(3))) double*;
    st = sched_lu_lds(s, lu, l, P, (LdsD)(double*)lds, (int)blockDim.x, (int)threadIdx.x, &d, &dp);
  } else if (compact == 2) {
    st = sched_lu_lds(s, lu, l, P, xg + (int64_t)blockIdx.x * xg_block, (int)blockDim.x, (int)threadIdx.x, &d, &dp);
  } else {
    st = sched_lu(s, lu, l, P, &d, &dp);
  }
  if (st == LU_OK) {
    lock[l] = 1;
  } else {
    dep[l] = d;
    dpiv[l] = dp;
  }
}

// dst[r * m + i] = src[r * stride + lanes[i]] for r < rows, i < m: rows of
// the listed lanes of a structure-of-arrays batch, packed.
__global__ void __launch_bounds__(256) k_gather_soa(const kcplx* src, int64_t stride, int64_t rows,
                                                    const int64_t* lanes, int64_t m, kcplx* dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * m) return;
  const int64_t r = t / m, i = t % m;
  dst[t] = src[r * stride + lanes[i]];
}

// The per-lane LU (lu_factor + lu_solve) of m packed lanes (lc, structure of
// arrays over i < m), lanes[i] being each one's lane in the chunk: its status
// into lustat[lanes[i]] and, on success, its solution into x[:, lanes[i]].
__global__ void __launch_bounds__(kBlock) k_ac_lane_lu(LuArgsT<kcplx> lc, const int64_t* lanes, int64_t m,
                                                       kcplx* x, int64_t BL, int32_t* lustat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int st = lu_factor(lc, i);
  if (st == LU_OK) st = lu_solve(lc, i);
  const int64_t l = lanes[i];
  lustat[l] = st;
  if (st == LU_OK)
    for (int k = 0; k < lc.n; k++) x[(int64_t)k * BL + l] = lc.x[(int64_t)k * m + i];
}

constexpr int kAcLaneBlock = 16;  // lanes per workgroup of k_ac_stamp and k_ac_solve on a chunk

// Per circuit, walks the chunk's frequencies in order as the sweep does: the
// first failing frequency sets the circuit's status and stops it.
__global__ void __launch_bounds__(256) k_ac_reduce(AcArgs a, int32_t* done, int32_t* status, int32_t* points,
                                                   double* maxrr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.B || done[c]) return;
  for (int64_t l = c * a.nfc; l < (c + 1) * a.nfc; l++) {
    const int s = a.lstatus[l];
    if (s < 0) { done[c] = 1; return; }  // the operating point failed: its status stands
    if (s == ST_OK || s == ST_LEDGER) maxrr[c] = fmax(maxrr[c], a.lrr[l]);
    if (s != ST_OK) {
      status[c] = s;
      done[c] = 1;
      return;
    }
    points[c]++;
  }
}

// ---- device buffers -----------------------------------------------------------------
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    release();
    if (b == 0) b = 8;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) { p = nullptr; return e; }
    bytes = b;
    return hipMemset(p, 0, b);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// LU workspace for B lanes of n unknowns with capL / capU entries, values of
// type T (double, or kcplx for AC systems).
template <typename T>
struct LuWorkT {
  DBuf xw, w, udiag, lx, ux, li, ui, lp, up, pinv, prow, xi, pstack, mark, err, rch, rp, dk, cpid;
  int64_t capL = 0, capU = 0, capR = 0, B = 0;
  int n = 0;
  hipError_t ensure(int n_, int64_t B_, int64_t cap) {
    if (n_ == n && B_ == B && cap <= capL && xw.p) return hipSuccess;
    n = n_;
    B = B_;
    capL = capU = cap;
    capR = capL + capU + n;
    size_t nd = (size_t)n * B * sizeof(T), ni = (size_t)n * B * 4;
    hipError_t e;
#define KH_A(buf, bytes_) if ((e = buf.alloc(bytes_)) != hipSuccess) return e
    KH_A(xw, nd); KH_A(w, nd); KH_A(udiag, nd);
    KH_A(lx, (size_t)capL * B * sizeof(T)); KH_A(ux, (size_t)capU * B * sizeof(T));
    KH_A(li, (size_t)capL * B * 4); KH_A(ui, (size_t)capU * B * 4);
    KH_A(lp, (size_t)(n + 1) * B * 4); KH_A(up, (size_t)(n + 1) * B * 4);
    KH_A(pinv, ni); KH_A(prow, ni); KH_A(xi, ni); KH_A(pstack, ni); KH_A(mark, ni);
    KH_A(err, (size_t)B * 4);
    KH_A(rch, (size_t)capR * B * 4); KH_A(rp, (size_t)(n + 1) * B * 4); KH_A(dk, ni);
    KH_A(cpid, (size_t)B * 4);
#undef KH_A
    return invalidate();
  }
  // Workspace of the dense backend (lu_dense): an n*n matrix in lx, udiag,
  // w, prow and err; nothing else is used.
  hipError_t ensure_dense(int n_, int64_t B_) {
    if (n_ == n && B_ == B && capL == (int64_t)n_ * n_ && capU == 0 && lx.p) return hipSuccess;
    n = n_;
    B = B_;
    capL = (int64_t)n * n;
    capU = capR = 0;
    size_t nd = (size_t)n * B * sizeof(T), ni = (size_t)n * B * 4;
    hipError_t e;
    if ((e = w.alloc(nd)) != hipSuccess) return e;
    if ((e = udiag.alloc(nd)) != hipSuccess) return e;
    if ((e = lx.alloc((size_t)capL * B * sizeof(T))) != hipSuccess) return e;
    if ((e = prow.alloc(ni)) != hipSuccess) return e;
    if ((e = err.alloc((size_t)B * 4)) != hipSuccess) return e;
    return hipSuccess;
  }
  // Forget the cached reach lists (a pattern id now names a different pattern).
  hipError_t invalidate() { return hipMemset(cpid.p, 0xFF, (size_t)B * 4); }  // -1: no cached reach
  void fill(LuArgsT<T>& a) const {
    a.capL = capL; a.capU = capU; a.capR = capR;
    a.rch = rch.as<int32_t>(); a.rp = rp.as<int32_t>(); a.dk = dk.as<int32_t>(); a.cpid = cpid.as<int32_t>();
    a.xw = xw.as<T>(); a.w = w.as<T>(); a.udiag = udiag.as<T>();
    a.lx = lx.as<T>(); a.ux = ux.as<T>(); a.li = li.as<int32_t>(); a.ui = ui.as<int32_t>();
    a.lp = lp.as<int32_t>(); a.up = up.as<int32_t>(); a.pinv = pinv.as<int32_t>(); a.prow = prow.as<int32_t>();
    a.xi = xi.as<int32_t>(); a.pstack = pstack.as<int32_t>(); a.mark = mark.as<int32_t>();
    a.err_step = err.as<int32_t>();
  }
};

using LuWork = LuWorkT<double>;

// L and U capacity (entries per system) for a pattern with structural fill
// max_fill = |L| + |U| + n under diagonal pivoting, with a wide margin for
// pivoting away from the diagonal. Batched analyses that still overflow it
// double it and refactor (WAIT_CAPACITY).

int64_t lu_capacity(int64_t max_fill, int n) { return 2 * max_fill + 4 * (int64_t)n + 64; }

// The batched analyses start L and U at the structural fill itself (each of
// L and U holds at most |L| + |U| + n entries under diagonal pivoting): a
// circuit's working set stays small (cache and LDS footprint), and the rare
// circuit whose pivoting fills more waits for the capacity to double.
int64_t lu_capacity_batch(int64_t max_fill, int n) { return max_fill + n + 16; }

// Pattern pool mirrored on the device.
struct DevPatterns {
  PatternPool host;
  DBuf desc, pool;
  size_t up_desc = 0, up_pool = 0;
  hipError_t sync() {
    if (host.desc.size() != up_desc || host.pool.size() != up_pool) {
      hipError_t e;
      if ((e = desc.alloc(host.desc.size() * sizeof(PatDesc))) != hipSuccess) return e;
      if ((e = pool.alloc(host.pool.size() * sizeof(int32_t))) != hipSuccess) return e;
      if ((e = hipMemcpy(desc.p, host.desc.data(), host.desc.size() * sizeof(PatDesc), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
      if ((e = hipMemcpy(pool.p, host.pool.data(), host.pool.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
      up_desc = host.desc.size();
      up_pool = host.pool.size();
    }
    return hipSuccess;
  }
};

// Compiled schedules (kh_sched.h) of the patterns of a pool, several per
// pattern (one per pivot sequence met), each uploaded to its own device
// buffer, with the device array of their views that k_ac_sched indexes.
struct DevScheds {
  struct One {
    DBuf buf;
    SchedView v;
    int64_t circuit = 0;  // the probe's circuit and frequency (where the schedule is likely to fit)
    int32_t freq = 0;
    uint64_t hash = 0;
    int64_t nl = 0, nu = 0;  // entries of L and U
  };
  DBuf lx, ux;  // the lanes' L and U values, [entries][lanes], grown to what a round's schedules need
  int64_t cap_lu = 0, cap_lanes = 0;
  DBuf xg;  // compact working columns in device memory (k_ac_sched compact 2)
  int64_t xg_doubles = 0;
  std::vector<std::unique_ptr<One>> all;
  DBuf dviews;
  size_t uploaded = 0;
  static uint64_t hash_of(const SchedHost& h) {  // FNV-1a over the structure
    uint64_t x = 1469598103934665603ull ^ (uint64_t)h.pid;
    auto mix = [&](const std::vector<int32_t>& v) {
      for (int32_t e : v) x = (x ^ (uint32_t)e) * 1099511628211ull;
      x = (x ^ v.size()) * 1099511628211ull;
    };
    mix(h.prow); mix(h.lp); mix(h.li); mix(h.up);
    return x;
  }
  // Index of the schedule with h's structure, adding it (uploaded) if new.
  hipError_t add(const SchedHost& h, const PatDesc& pd, const int32_t* dpool, int64_t circuit, int32_t freq,
                 int32_t* index) {
    const uint64_t hh = hash_of(h);
    for (size_t i = 0; i < all.size(); i++)
      if (all[i]->hash == hh && all[i]->v.pid == h.pid) {
        *index = (int32_t)i;
        return hipSuccess;
      }
    auto o = std::make_unique<One>();
    std::vector<int32_t> pk;
    auto put = [&](const std::vector<int32_t>& v) {
      const size_t at = pk.size();
      pk.insert(pk.end(), v.begin(), v.end());
      while (pk.size() % 4) pk.push_back(0);  // 16-byte alignment of every array
      return at;
    };
    const size_t o_prow = put(h.prow), o_lp = put(h.lp), o_li = put(h.li), o_up = put(h.up), o_dc = put(h.dc),
                 o_tb = put(h.tbase), o_tp = put(h.tpos), o_ap = put(h.apos), o_ds = put(h.dslot);
    const size_t o_ue = pk.size();
    pk.resize(o_ue + 4 * h.ue.size());
    if (!h.ue.empty()) std::memcpy(pk.data() + o_ue, h.ue.data(), h.ue.size() * sizeof(SchedU));
    hipError_t e = o->buf.alloc(pk.size() * 4);
    if (e != hipSuccess) return e;
    e = hipMemcpy(o->buf.p, pk.data(), pk.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const int32_t* d = o->buf.as<int32_t>();
    o->v.n = h.n;
    o->v.pid = h.pid;
    o->v.prow = (SchedPtr<int32_t>)(d + o_prow);
    o->v.lp = (SchedPtr<int32_t>)(d + o_lp);
    o->v.li = (SchedPtr<int32_t>)(d + o_li);
    o->v.up = (SchedPtr<int32_t>)(d + o_up);
    o->v.dc = (SchedPtr<int32_t>)(d + o_dc);
    o->v.mmax = h.mmax;
    o->v.tbase = (SchedPtr<int32_t>)(d + o_tb);
    o->v.tpos = (SchedPtr<int32_t>)(d + o_tp);
    o->v.apos = (SchedPtr<int32_t>)(d + o_ap);
    o->v.dslot = (SchedPtr<int32_t>)(d + o_ds);
    o->v.ue = (SchedPtr<SchedU>)(d + o_ue);
    o->v.ap = (SchedPtr<int32_t>)(dpool + pd.ap);
    o->v.arow = (SchedPtr<int32_t>)(dpool + pd.arow);
    o->v.aslot = (SchedPtr<int32_t>)(dpool + pd.aslot);
    o->v.q = (SchedPtr<int32_t>)(dpool + pd.q);
    o->v.qinv = (SchedPtr<int32_t>)(dpool + pd.qinv);
    o->circuit = circuit;
    o->freq = freq;
    o->hash = hh;
    o->nl = (int64_t)h.li.size();
    o->nu = (int64_t)h.ue.size();
    *index = (int32_t)all.size();
    all.push_back(std::move(o));
    return hipSuccess;
  }
  hipError_t sync() {  // the views array k_ac_sched reads
    if (uploaded == all.size()) return hipSuccess;
    std::vector<SchedView> v;
    for (const auto& o : all) v.push_back(o->v);
    hipError_t e = dviews.alloc(v.size() * sizeof(SchedView));
    if (e == hipSuccess) e = hipMemcpy(dviews.p, v.data(), v.size() * sizeof(SchedView), hipMemcpyHostToDevice);
    if (e == hipSuccess) uploaded = all.size();
    return e;
  }
};

// Lanes per workgroup of k_ac_sched: 16 (a complex128 access of 16 lanes is
// two whole cache lines), fewer when the chunk would otherwise not give every
// SIMD a wavefront (the lockstep chain is latency-bound: wavefronts in flight
// beat fuller ones) or when the compact working columns of mmax slots per
// lane would not fit kSchedLds, at least 4; KH_SCHED_BLOCK overrides it.
// LDS a workgroup of k_ac_sched may hold for its compact working columns.
constexpr int64_t kSchedLds = 40 * 1024;

int sched_block(int64_t lanes, int64_t mmax) {
  if (const char* e = std::getenv("KH_SCHED_BLOCK")) {
    const int b = std::atoi(e);
    if (b >= 1 && b <= 64 && (b & (b - 1)) == 0) return b;
  }
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    ncu = 256;
  // the widest block whose compact working columns fit the LDS budget
  int fit = 16;
  while (fit > 4 && mmax * fit * (int64_t)sizeof(kcplx) > kSchedLds) fit /= 2;
  for (int b : {16, 8})
    if (b <= fit && lanes >= (int64_t)b * 4 * ncu) return b;
  return std::min(4, fit);
}

// Gathers lanes[k]'s nslots values (structure-of-arrays [nslots][stride]) into
// dst[k][nslots].
__global__ void __launch_bounds__(256) k_gather_lanes(const kcplx* src, int64_t stride, int64_t nslots,
                                                      const int64_t* lanes, int32_t nl, kcplx* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots * nl) return;
  const int64_t k = i / nslots, s = i % nslots;
  dst[i] = src[s * stride + lanes[k]];
}

// One chunk of AC lanes through compiled schedules.
//  * Probes: systems factored on the host (in parallel, kh_sched.h
//    build_schedule) into schedules; identical structures are kept once.
//  * Rounds: every workgroup with lanes left is given the untried schedule
//    whose probe is nearest (same circuit first, then frequency) and runs them
//    in lockstep. Before round 0 each circuit's frequencies are bisected with
//    probes (pivot sequences change in bands of frequency); a later round's
//    probes are one member of each run of lanes that departed together (same
//    circuit, schedule, column and pivot row, consecutive frequencies).
//  * Solo round: once at most kSolo lanes are left, each is its own probe and
//    runs alone (one lane per workgroup), all in one pass.
// lock[l] = 1 for the lanes solved; `left` receives the lanes with an
// operating point that were not (the per-lane LU's).
struct SchedChunk {
  int64_t ok = 0, left = 0, probes = 0, rounds = 0;
  double kernel_ms = 0, host_s = 0;
};

kh_status sched_chunk(const AcArgs& a, const LuArgsT<kcplx>& lu, DevScheds& S, const DevPatterns& pats,
                      const std::vector<int32_t>& lp, const std::vector<int32_t>& wst, int n, int nslots, int64_t cap,
                      int32_t* d_lock, int32_t* d_dep, int32_t* d_dpiv, hipEvent_t ev[2], SchedChunk& out,
                      std::vector<int64_t>& left) {
  constexpr int kRounds = 4, kFirstProbes = 9, kMaxProbes = 1024, kSolo = 4096;
  constexpr int64_t kBisect = 8;
  int64_t lds_total = 0;  // LDS of the whole device
  {
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess)
      lds_total = (int64_t)ncu * per_cu;
  }
  const int64_t BL = a.BL, nfc = a.nfc;
  std::vector<int32_t> hl(BL, 0), hd(BL, -1), hp(BL, -1);
  std::vector<int8_t> refused(BL, 0);  // a check other than the pivot rule failed: per-lane LU
  KH_HIP(hipMemsetAsync(d_lock, 0, (size_t)BL * 4, 0));
  auto open = [&](int64_t l) { return wst[l / nfc] == ST_OK && !hl[l] && !refused[l]; };
  // Factors the probe lanes on the host; idx[i] = schedule of probes[i] (-1: did not factor).
  auto factor_probes = [&](const std::vector<int64_t>& probes, std::vector<int32_t>& idx) -> kh_status {
    idx.assign(probes.size(), -1);
    if (probes.empty()) return KH_OK;
    const auto th0 = std::chrono::steady_clock::now();
    const int np = (int)probes.size();
    DBuf dl, dv;
    KH_HIP(dl.alloc((size_t)np * 8));
    KH_HIP(dv.alloc((size_t)np * nslots * sizeof(kcplx)));
    KH_HIP(hipMemcpy(dl.p, probes.data(), (size_t)np * 8, hipMemcpyHostToDevice));
    const int64_t tot = (int64_t)np * nslots;
    hipLaunchKernelGGL(k_gather_lanes, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, a.val, BL,
                       (int64_t)nslots, dl.as<int64_t>(), np, dv.as<kcplx>());
    KH_HIP(hipGetLastError());
    std::vector<kcplx> pv((size_t)tot);
    KH_HIP(hipMemcpy(pv.data(), dv.p, (size_t)tot * sizeof(kcplx), hipMemcpyDeviceToHost));
    std::vector<SchedHost> hs(np);
    std::vector<char> built(np, 0);
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int i; (i = next++) < np;)
        built[i] = build_schedule<kcplx>(pats.host.desc, pats.host.pool, lp[probes[i]], n, nslots,
                                         pv.data() + (size_t)i * nslots, cap, hs[i]);
    };
    const int nt = std::max(1, std::min<int>(np, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    for (int i = 0; i < np; i++) {
      if (!built[i]) {
        refused[probes[i]] = 1;  // the probe itself does not factor: it goes to the per-lane LU
        continue;
      }
      KH_HIP(S.add(hs[i], pats.host.desc[lp[probes[i]]], pats.pool.as<int32_t>(), probes[i] / nfc,
                   a.f0 + (int32_t)(probes[i] % nfc), &idx[i]));
    }
    KH_HIP(S.sync());
    out.probes += np;
    out.host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - th0).count();
    return KH_OK;
  };
  // Runs workgroups of W lanes with schedules wgv (-1: none), then reads back
  // which lanes were solved and where the others departed.
  auto run_round = [&](int W, const std::vector<int32_t>& wgv) -> kh_status {
    // L and U values for the largest schedule of the round (each lane's
    // factors live only within the launch that factors and solves it).
    int64_t need = 1;
    for (int32_t v : wgv)
      if (v >= 0) need = std::max({need, S.all[v]->nl, S.all[v]->nu});
    if (need > S.cap_lu || BL > S.cap_lanes) {
      S.cap_lu = std::max(need, S.cap_lu);
      S.cap_lanes = std::max((BL + 63) / 64 * 64, S.cap_lanes);  // whole blocks of up to 64 lanes
      KH_HIP(S.lx.alloc((size_t)S.cap_lu * S.cap_lanes * sizeof(kcplx)));
      KH_HIP(S.ux.alloc((size_t)S.cap_lu * S.cap_lanes * sizeof(kcplx)));
    }
    LuArgsT<kcplx> lr = lu;
    lr.lx = S.lx.as<kcplx>();
    lr.ux = S.ux.as<kcplx>();
    lr.capL = lr.capU = S.cap_lu;
    // Compact working columns in LDS when the round's largest column fits.
    int64_t mm = 1;
    for (int32_t v : wgv)
      if (v >= 0) mm = std::max<int64_t>(mm, S.all[v]->v.mmax);
    const int64_t lds = mm * W * (int64_t)sizeof(kcplx);
    // LDS only when every active workgroup of the round fits at once (LDS
    // capacity bounds the workgroups resident per compute unit, and a second
    // pass of the latency-bound chain costs more than LDS saves).
    int64_t active = 0;
    for (int32_t v : wgv) active += v >= 0;
    // Otherwise compact in device memory: a workgroup's block of mm slots is
    // small enough to stay in the caches, where the dense working columns
    // (n slots per lane) spill to HBM.
    const int compact = (lds <= kSchedLds && active * lds <= lds_total) ? 1 : 2;
    const int64_t xg_block = mm * W * (int64_t)(sizeof(kcplx) / sizeof(double));
    if (compact == 2 && (int64_t)wgv.size() * xg_block > S.xg_doubles) {
      S.xg_doubles = (int64_t)wgv.size() * xg_block;
      KH_HIP(S.xg.alloc((size_t)S.xg_doubles * sizeof(double)));
    }
    DBuf dwgv;
    KH_HIP(dwgv.alloc(wgv.size() * 4));
    KH_HIP(hipMemcpy(dwgv.p, wgv.data(), wgv.size() * 4, hipMemcpyHostToDevice));
    KH_HIP(hipEventRecord(ev[0], 0));
    hipLaunchKernelGGL(k_ac_sched, dim3((unsigned)wgv.size()), dim3(W), compact == 1 ? (size_t)lds : 0, 0, a, lr,
                       S.dviews.as<SchedView>(), dwgv.as<int32_t>(), d_lock, d_dep, d_dpiv, compact,
                       compact == 2 ? S.xg.as<double>() : nullptr, xg_block);
    KH_HIP(hipGetLastError());
    KH_HIP(hipEventRecord(ev[1], 0));
    KH_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    KH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    out.kernel_ms += ms;
    out.rounds++;
    KH_HIP(hipMemcpy(hl.data(), d_lock, (size_t)BL * 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hd.data(), d_dep, (size_t)BL * 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hp.data(), d_dpiv, (size_t)BL * 4, hipMemcpyDeviceToHost));
    if (std::getenv("KH_SCHED_LOG")) {  // KH_SCHED_LOG=1: one line per round on stderr
      int64_t ok = 0, open_n = 0, ref = 0;
      for (int64_t l = 0; l < BL; l++) {
        ok += hl[l] != 0;
        ref += refused[l] != 0;
        open_n += open(l);
      }
      fprintf(stderr,
              "kh_sched: round %lld W %d: %.1f ms, solved %lld, open %lld, refused %lld, schedules %zu, L/U %lld, "
              "slots %lld (%s)\n",
              (long long)out.rounds, W, ms, (long long)ok, (long long)open_n, (long long)ref, S.all.size(),
              (long long)need, (long long)mm, compact == 1 ? "LDS" : "compact in device memory");
    }
    return KH_OK;
  };
  int W = 0;  // lanes per workgroup, set once round 0's schedules are known
  int64_t nwg = 0;
  std::vector<std::vector<int32_t>> tried;
  // Every probe of this chunk: (circuit, frequency, schedule). Identical
  // structures share one schedule, so a schedule may have several probes.
  struct ProbeAt {
    int64_t circuit;
    int32_t freq, sched;
  };
  std::vector<ProbeAt> at;
  // Round 0's probes: pivot sequences change in bands of frequency, so each
  // circuit's frequencies are bisected — kFirstProbes evenly spaced lanes
  // (its first and last included), then the middle of every interval whose
  // two ends got different schedules, until such intervals are at most
  // kBisect lanes wide. Later rounds catch bands no probe landed in.
  std::vector<int64_t> probes;
  {
    std::map<int64_t, int32_t> got;  // probed lane -> schedule (-1: did not factor)
    for (int64_t c = 0; c < a.B; c++) {
      if (wst[c] != ST_OK) continue;
      const int k0 = (int)std::min<int64_t>(kFirstProbes, nfc);
      for (int k = 0; k < k0; k++) probes.push_back(c * nfc + (k0 > 1 ? k * (nfc - 1) / (k0 - 1) : 0));
    }
    std::sort(probes.begin(), probes.end());
    probes.erase(std::unique(probes.begin(), probes.end()), probes.end());
    while (!probes.empty()) {
      if ((int64_t)probes.size() > kMaxProbes) probes.resize(kMaxProbes);
      std::vector<int32_t> idx;
      if (kh_status e = factor_probes(probes, idx); e != KH_OK) return e;
      for (size_t i = 0; i < probes.size(); i++) {
        got[probes[i]] = idx[i];
        if (idx[i] >= 0) at.push_back({probes[i] / nfc, a.f0 + (int32_t)(probes[i] % nfc), idx[i]});
      }
      probes.clear();
      for (auto it = got.begin(); it != got.end(); ++it) {
        auto nx = std::next(it);
        if (nx == got.end() || nx->first / nfc != it->first / nfc) continue;
        if (nx->first - it->first > kBisect && it->second != nx->second)
          probes.push_back((it->first + nx->first) / 2);
      }
    }
  }
  for (int round = 0; round < kRounds; round++) {
    if (round > 0) {
      // Few lanes left: the solo round below serves them in one pass.
      int64_t left_n = 0;
      for (int64_t l = 0; l < BL; l++) left_n += open(l);
      if (left_n <= kSolo) break;
    }
    if ((int64_t)probes.size() > kMaxProbes) probes.resize(kMaxProbes);
    std::vector<int32_t> idx;
    if (kh_status e = factor_probes(probes, idx); e != KH_OK) return e;
    for (size_t i = 0; i < probes.size(); i++)
      if (idx[i] >= 0) at.push_back({probes[i] / nfc, a.f0 + (int32_t)(probes[i] % nfc), idx[i]});
    if (W == 0) {
      int64_t mm = 1;
      for (const auto& o : S.all) mm = std::max<int64_t>(mm, o->v.mmax);
      W = sched_block(BL, mm);
      nwg = (BL + W - 1) / W;
      tried.assign(nwg, {});
    }
    std::vector<int32_t> wgv(nwg, -1);
    bool any = false;
    for (int64_t w = 0; w < nwg; w++) {
      int64_t l0 = -1;
      for (int64_t l = w * W; l < std::min(BL, (w + 1) * W); l++)
        if (open(l)) {
          l0 = l;
          break;
        }
      if (l0 < 0) continue;
      const int32_t pid = lp[l0];
      const int64_t c = l0 / nfc;
      const int32_t f = a.f0 + (int32_t)(l0 % nfc);
      int32_t best = -1;
      int64_t bestd = 0;
      auto consider = [&](int32_t i, int64_t circuit, int32_t freq) {
        if (S.all[i]->v.pid != pid) return;
        if (std::find(tried[w].begin(), tried[w].end(), i) != tried[w].end()) return;
        const int64_t d = (circuit == c ? 0 : (int64_t)1 << 40) + std::llabs((int64_t)freq - f);
        if (best < 0 || d < bestd) {
          best = i;
          bestd = d;
        }
      };
      for (const ProbeAt& q : at) consider(q.sched, q.circuit, q.freq);
      for (size_t i = 0; i < S.all.size(); i++) consider((int32_t)i, S.all[i]->circuit, S.all[i]->freq);
      if (best < 0) continue;
      wgv[w] = best;
      tried[w].push_back(best);
      any = true;
    }
    if (!any) break;
    if (kh_status e = run_round(W, wgv); e != KH_OK) return e;
    // Next probes: runs of consecutive lanes that departed together.
    probes.clear();
    std::vector<int64_t> run;
    auto flush = [&]() {
      if (!run.empty()) probes.push_back(run[run.size() / 2]);
      run.clear();
    };
    int64_t prev = -2;
    std::tuple<int32_t, int32_t, int32_t> prev_key{-1, -1, -1};
    for (int64_t w = 0; w < nwg; w++) {
      if (wgv[w] < 0) continue;
      for (int64_t l = w * W; l < std::min(BL, (w + 1) * W); l++) {
        if (!open(l) || lp[l] != S.all[wgv[w]]->v.pid) continue;
        if (hd[l] < 0) {
          refused[l] = 1;
          continue;
        }
        const auto key = std::make_tuple(wgv[w], hd[l], hp[l]);
        if (l != prev + 1 || l / nfc != prev / nfc || key != prev_key) flush();
        run.push_back(l);
        prev = l;
        prev_key = key;
      }
    }
    flush();
  }
  // Solo round: every lane left is its own probe.
  std::vector<int64_t> rest;
  for (int64_t l = 0; l < BL; l++)
    if (open(l)) rest.push_back(l);
  if (!rest.empty() && (int64_t)rest.size() <= kSolo) {
    std::vector<int32_t> idx;
    if (kh_status e = factor_probes(rest, idx); e != KH_OK) return e;
    std::vector<int32_t> wgv(BL, -1);
    for (size_t i = 0; i < rest.size(); i++) wgv[rest[i]] = idx[i];
    if (kh_status e = run_round(1, wgv); e != KH_OK) return e;
  }
  for (int64_t l = 0; l < BL; l++) {
    if (wst[l / nfc] != ST_OK) continue;
    if (hl[l]) {
      out.ok++;
    } else {
      out.left++;
      left.push_back(l);
    }
  }
  return KH_OK;
}

std::string slots_key(const std::vector<int32_t>& slots) {
  return std::string(reinterpret_cast<const char*>(slots.data()), slots.size() * sizeof(int32_t));
}

inline unsigned grid_for(int64_t B) { return (unsigned)((B + kBlock - 1) / kBlock); }

}  // namespace

// ---- handles -------------------------------------------------------------------------
struct kh_topology_s {
  Topology t;
  DBuf et, isV, slot_i, slot_j, diag, rowmajor;  // rowmajor: slots sorted by (row, column)
  DBuf lists;  // nonlinear, reactive and source element indices, concatenated
  // the wavefront form's stamping plan and AbsProduct rows (WavePlan), concatenated:
  // ch_el, ch_t, tgt, tel_ptr, tel, rec_off, ro_ptr / ro_slot sparse, ro_ptr / ro_slot dense
  DBuf wave;
  int64_t wave_off[10] = {};
  int32_t nchunks = 0, rec_cap = 0;
  std::mutex mu;  // batch entry points on one topology run one at a time
};

struct kh_lu_batch_s {
  int n = 0;
  int64_t nnz = 0, B = 0;
  DevPatterns pats;
  Pattern pat;
  LuWork work;
  DBuf pid, status;
};

// The per-system seam (solver.Sparse / CSparse behind Matrix / CMatrix): the
// stamps of one system in first-Add order (sparseLU.add, solver/sparse.go:
// 114-143), its pattern and device workspace. T = double or kcplx.
template <typename T>
struct SysT {
  int n = 0;
  std::unordered_map<int64_t, int32_t> slot;
  std::vector<int32_t> ri, ci;
  std::vector<T> val;
  bool dirty = true;         // the pattern changed since the last analysis
  bool valid = false;
  uint64_t generation = 0;  // counts pattern analyses (the static scratch's owner tag)
  uint64_t dev_generation = 0;  // the analysis the device buffers were sized for (0: none)
  Pattern pat;
  DevPatterns pats;
  LuWorkT<T> work;
  DBuf dval, drhs, dx, pid, status;
  size_t dval_slots = 0;
};
struct kh_sys_s : SysT<double> {};
struct kh_csys_s : SysT<kcplx> {};

extern "C" {

const char* kh_status_string(kh_status s) {
  switch (s) {
    case KH_OK: return "ok";
    case KH_SINGULAR: return "solver: singular matrix";
    case KH_NONFINITE: return "solver: non-finite value in system";
    case KH_BADSIZE: return "solver: invalid matrix size";
    case KH_NOCONV: return "analysis: no convergence";
    case KH_TSTEP_TOO_SMALL: return "analysis: timestep too small";
    case KH_TOO_MANY_STEPS: return "analysis: too many time steps";
    case KH_INTERRUPTED: return "analysis interrupted";
    case KH_LEDGER: return "em: energy ledger does not balance";
    case KH_BAD_TRAN_SPEC: return "analysis: invalid transient specification";
    case KH_UNSUPPORTED: return "device: unsupported element kind";
    case KH_BAD_ARGUMENT: return "invalid argument";
    case KH_HIP_ERROR: return "HIP runtime error";
    case KH_NO_DEVICE: return "no gfx950 device";
    case KH_CAPACITY: return "L/U capacity exceeded";
    case KH_RUNNING: return "running";
  }
  return "unknown status";
}

const char* kh_last_error(void) { return g_last_error.c_str(); }

kh_status kh_init(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(KH_NO_DEVICE, "no HIP device");
  int dev = 0;
  KH_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  KH_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(KH_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", the library is built for gfx950");
  return KH_OK;
}

// ---- per-system seam --------------------------------------------------------------------
}  // extern "C"

namespace {

template <typename T>
kh_status sys_add(SysT<T>* s, const int32_t* i, const int32_t* j, const T* v, int64_t count) {
  for (int64_t k = 0; k < count; k++) {  // sparseLU.add, solver/sparse.go:114-143
    int ii = i[k], jj = j[k];
    if (ii < 0 || jj < 0 || ii >= s->n || jj >= s->n) continue;
    int64_t key = (int64_t)ii * s->n + jj;
    auto it = s->slot.find(key);
    int32_t sl;
    if (it == s->slot.end()) {
      sl = (int32_t)s->val.size();
      s->slot[key] = sl;
      s->ri.push_back(ii);
      s->ci.push_back(jj);
      s->val.push_back(lu_zero_value<T>());
      s->dirty = true;
    } else {
      sl = it->second;
    }
    s->val[sl] = s->val[sl] + v[k];
    s->valid = false;
  }
  return KH_OK;
}

// Symbolic phase of a grown pattern (analyze, solver/sparse.go:154-187).
template <typename T>
void sys_analyze(SysT<T>* s) {
  if (!s->dirty) return;
  std::vector<int32_t> slots(s->val.size());
  for (size_t k = 0; k < slots.size(); k++) slots[k] = (int32_t)k;
  s->pat = analyze_pattern(s->n, s->ri, s->ci, slots);
  s->pats.host = PatternPool();
  s->pats.up_desc = s->pats.up_pool = 0;
  s->pats.host.add(s->pat, slots_key(slots));
  s->generation++;
  s->dirty = false;
}

template <typename T>
kh_status sys_prepare(SysT<T>* s) {
  sys_analyze(s);
  if (s->dev_generation != s->generation) {
    KH_HIP(s->pats.sync());
    KH_HIP(s->work.ensure(s->n, 1, lu_capacity(s->pat.fill_estimate, s->n)));
    KH_HIP(s->work.invalidate());  // pattern id 0 is reused for the new pattern
    KH_HIP(s->dval.alloc(std::max<size_t>(1, s->val.size()) * sizeof(T)));
    s->dval_slots = s->val.size();
    KH_HIP(s->drhs.alloc((size_t)s->n * sizeof(T)));
    KH_HIP(s->dx.alloc((size_t)s->n * sizeof(T)));
    KH_HIP(s->pid.alloc(4));
    KH_HIP(s->status.alloc(4));
    s->dev_generation = s->generation;
  }
  return KH_OK;
}

// The factor + solve of one system (B = 1) on one lane: the seams
// (kh_sys_solve, kh_csys_solve, kh_sys_solve_static).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_lu_t(LuArgsT<T> a, int32_t* status) {
  if (threadIdx.x != 0) return;
  int st = lu_factor(a, 0);
  if (st == LU_OK) st = lu_solve(a, 0);
  status[0] = st;
}
template <typename T>
hipError_t launch_sys(const LuArgsT<T>& a, int32_t* status, hipStream_t st) {
  hipLaunchKernelGGL(k_lu_t<T>, dim3(1), dim3(kBlock), 0, st, a, status);
  return hipGetLastError();
}

kh_status lu_status_error(int32_t hs) {
  if (hs == LU_SINGULAR) return fail(KH_SINGULAR, "solver: singular matrix");
  if (hs == LU_NONFINITE) return fail(KH_NONFINITE, "solver: non-finite value in system");
  if (hs == LU_CAPACITY) return fail(KH_CAPACITY, "L/U capacity exceeded");
  return KH_OK;
}

template <typename T>
kh_status sys_solve(SysT<T>* s, const T* b, T* x) {
  kh_status st = sys_prepare(s);
  if (st != KH_OK) return st;
  if (s->val.empty()) return fail(KH_SINGULAR, "empty matrix");
  KH_HIP(hipMemcpy(s->dval.p, s->val.data(), s->val.size() * sizeof(T), hipMemcpyHostToDevice));
  KH_HIP(hipMemcpy(s->drhs.p, b, (size_t)s->n * sizeof(T), hipMemcpyHostToDevice));
  LuArgsT<T> a = {};
  a.n = s->n;
  a.nslots = (int32_t)s->val.size();
  a.B = 1;
  a.pats = s->pats.desc.template as<PatDesc>();
  a.pool = s->pats.pool.template as<int32_t>();
  a.pid = s->pid.template as<int32_t>();
  a.val = s->dval.template as<T>();
  a.rhs = s->drhs.template as<T>();
  a.x = s->dx.template as<T>();
  s->work.fill(a);
  KH_HIP(launch_sys(a, s->status.template as<int32_t>(), 0));
  int32_t hs = 0;
  KH_HIP(hipMemcpy(&hs, s->status.p, 4, hipMemcpyDeviceToHost));
  if (hs == LU_OK || hs == LU_NONFINITE) KH_HIP(hipMemcpy(x, s->dx.p, (size_t)s->n * sizeof(T), hipMemcpyDeviceToHost));
  s->valid = hs == LU_OK;
  return lu_status_error(hs);
}

// AbsProduct, solver/sparse.go:481-490 (real) and :519-527 (complex): y_i =
// sum over the stored entries, in storage (first-Add) order, of |A_ij|·|x_j|.
template <typename T>
void sys_absproduct(const SysT<T>* s, const T* x, double* y) {
  for (int i = 0; i < s->n; i++) y[i] = 0;
  for (size_t k = 0; k < s->val.size(); k++) y[s->ri[k]] += lu_mag(s->val[k]) * lu_mag(x[s->ci[k]]);
}

template <typename T>
kh_status sys_nnz(const SysT<T>* s, int64_t* a, int64_t* l, int64_t* u) {
  if (!s->valid) return fail(KH_BAD_ARGUMENT, "no factorization held");
  int32_t lpn = 0, upn = 0;
  KH_HIP(hipMemcpy(&lpn, s->work.lp.template as<int32_t>() + s->n, 4, hipMemcpyDeviceToHost));
  KH_HIP(hipMemcpy(&upn, s->work.up.template as<int32_t>() + s->n, 4, hipMemcpyDeviceToHost));
  if (a) *a = (int64_t)s->val.size();
  if (l) *l = lpn;
  if (u) *u = (int64_t)upn + s->n;
  return KH_OK;
}

template <typename T>
kh_status sys_pivots(const SysT<T>* s, int32_t* q, int32_t* prow) {
  if (!s->valid) return fail(KH_BAD_ARGUMENT, "no factorization held");
  if (q) std::memcpy(q, s->pat.q.data(), (size_t)s->n * 4);
  if (prow) KH_HIP(hipMemcpy(prow, s->work.prow.p, (size_t)s->n * 4, hipMemcpyDeviceToHost));
  return KH_OK;
}

// ---- the allocation-free seam: kh_sys_solve_static --------------------------------------
// One system's pattern, values, solution and LU workspace in a file-scope
// device array (per device: each GPU holds its own copy, at its own address,
// until the process ends). Calls on one GPU share it and race if they
// overlap; the array keeps the last call's contents, so every call writes
// what the kernels read before they write it (the pattern when another
// system or pattern used the array last, the reach-cache tag).
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];

struct ScratchDev {
  char* ptr = nullptr;          // the device address of `scratch` on this device
  const void* owner = nullptr;  // the system whose pattern the array holds
  uint64_t generation = 0;
};
std::mutex g_scratch_mu;
ScratchDev g_scratch[kMaxDevices];

// The current device's copy of `scratch` (hipGetSymbolAddress once per device).
kh_status scratch_dev(ScratchDev** out) {
  int dev = 0;
  KH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(KH_NO_DEVICE, "device index out of range");
  ScratchDev& d = g_scratch[dev];
  if (!d.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(KH_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.ptr = static_cast<char*>(p);
  }
  *out = &d;
  return KH_OK;
}

// Byte layout of one system in the scratch array.
struct ScratchLayout {
  size_t desc, pool, pid, cpid, val, rhs, x, status, xw, w, udiag, lx, ux, li, ui, lp, up, pinv, prow, xi, pstack,
      mark, err, rch, rp, dk, total;
  int64_t capL = 0, capR = 0;
  void plan(int n, int64_t nslots, size_t pool_words, int64_t cap) {
    capL = cap;
    capR = 2 * cap + n;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t at = off;
      off += (bytes + 255) & ~(size_t)255;
      return at;
    };
    desc = take(sizeof(PatDesc));
    pool = take(pool_words * 4);
    pid = take(4);
    cpid = take(4);
    status = take(4);
    err = take(4);
    val = take((size_t)nslots * 8);
    rhs = take((size_t)n * 8);
    x = take((size_t)n * 8);
    xw = take((size_t)n * 8);
    w = take((size_t)n * 8);
    udiag = take((size_t)n * 8);
    lx = take((size_t)cap * 8);
    ux = take((size_t)cap * 8);
    li = take((size_t)cap * 4);
    ui = take((size_t)cap * 4);
    lp = take((size_t)(n + 1) * 4);
    up = take((size_t)(n + 1) * 4);
    pinv = take((size_t)n * 4);
    prow = take((size_t)n * 4);
    xi = take((size_t)n * 4);
    pstack = take((size_t)n * 4);
    mark = take((size_t)n * 4);
    rch = take((size_t)capR * 4);
    rp = take((size_t)(n + 1) * 4);
    dk = take((size_t)n * 4);
    total = off;
  }
};

}  // namespace

extern "C" {

kh_status kh_sys_create(int32_t n, kh_sys* out) {
  if (!out) return fail(KH_BAD_ARGUMENT, "null handle");
  if (n <= 0 || n > (1 << 22)) return fail(KH_BADSIZE, "size " + std::to_string(n));
  auto* s = new kh_sys_s();
  s->n = n;
  *out = s;
  return KH_OK;
}

kh_status kh_sys_clear(kh_sys s) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  std::fill(s->val.begin(), s->val.end(), 0.0);
  s->valid = false;
  return KH_OK;
}

kh_status kh_sys_add_batch(kh_sys s, const int32_t* i, const int32_t* j, const double* v, int64_t count) {
  if (!s || (count > 0 && (!i || !j || !v))) return fail(KH_BAD_ARGUMENT, "null argument");
  return sys_add<double>(s, i, j, v, count);
}

kh_status kh_sys_solve(kh_sys s, const double* b, double* x) {
  if (!s || !b || !x) return fail(KH_BAD_ARGUMENT, "null argument");
  return sys_solve<double>(s, b, x);
}

kh_status kh_sys_solve_static(kh_sys s, const double* b, double* x) {
  if (!s || !b || !x) return fail(KH_BAD_ARGUMENT, "null argument");
  if (s->val.empty()) return fail(KH_SINGULAR, "empty matrix");
  sys_analyze(s);  // host only: the device buffers of kh_sys_solve stay as they are
  const std::vector<int32_t>& pool = s->pats.host.pool;
  ScratchLayout L;
  L.plan(s->n, (int64_t)s->val.size(), pool.size(), lu_capacity(s->pat.fill_estimate, s->n));
  if (L.total > kScratchBytes)
    return fail(KH_BADSIZE, "kh_sys_solve_static: the system needs " + std::to_string(L.total) +
                                " bytes of workspace, above the static " + std::to_string(kScratchBytes));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchDev* d = nullptr;
  kh_status st = scratch_dev(&d);
  if (st != KH_OK) return st;
  char* base = d->ptr;
  if (d->owner != s || d->generation != s->generation) {
    // Another system (or pattern) used the array last: its pattern, the
    // pattern id and an empty reach cache.
    const int32_t zero = 0, none = -1;
    KH_HIP(hipMemcpy(base + L.desc, s->pats.host.desc.data(), sizeof(PatDesc), hipMemcpyHostToDevice));
    KH_HIP(hipMemcpy(base + L.pool, pool.data(), pool.size() * 4, hipMemcpyHostToDevice));
    KH_HIP(hipMemcpy(base + L.pid, &zero, 4, hipMemcpyHostToDevice));
    KH_HIP(hipMemcpy(base + L.cpid, &none, 4, hipMemcpyHostToDevice));
    d->owner = s;
    d->generation = s->generation;
  }
  KH_HIP(hipMemcpy(base + L.val, s->val.data(), s->val.size() * 8, hipMemcpyHostToDevice));
  KH_HIP(hipMemcpy(base + L.rhs, b, (size_t)s->n * 8, hipMemcpyHostToDevice));
  LuArgs a = {};
  a.n = s->n;
  a.nslots = (int32_t)s->val.size();
  a.B = 1;
  a.capL = a.capU = L.capL;
  a.capR = L.capR;
  a.pats = (const PatDesc*)(base + L.desc);
  a.pool = (const int32_t*)(base + L.pool);
  a.pid = (const int32_t*)(base + L.pid);
  a.val = (const double*)(base + L.val);
  a.rhs = (const double*)(base + L.rhs);
  a.x = (double*)(base + L.x);
  a.xw = (double*)(base + L.xw);
  a.w = (double*)(base + L.w);
  a.udiag = (double*)(base + L.udiag);
  a.lx = (double*)(base + L.lx);
  a.ux = (double*)(base + L.ux);
  a.li = (int32_t*)(base + L.li);
  a.ui = (int32_t*)(base + L.ui);
  a.lp = (int32_t*)(base + L.lp);
  a.up = (int32_t*)(base + L.up);
  a.pinv = (int32_t*)(base + L.pinv);
  a.prow = (int32_t*)(base + L.prow);
  a.xi = (int32_t*)(base + L.xi);
  a.pstack = (int32_t*)(base + L.pstack);
  a.mark = (int32_t*)(base + L.mark);
  a.err_step = (int32_t*)(base + L.err);
  a.rch = (int32_t*)(base + L.rch);
  a.rp = (int32_t*)(base + L.rp);
  a.dk = (int32_t*)(base + L.dk);
  a.cpid = (int32_t*)(base + L.cpid);
  KH_HIP(launch_sys(a, (int32_t*)(base + L.status), 0));
  int32_t hs = 0;
  KH_HIP(hipMemcpy(&hs, base + L.status, 4, hipMemcpyDeviceToHost));
  if (hs == LU_OK || hs == LU_NONFINITE) KH_HIP(hipMemcpy(x, base + L.x, (size_t)s->n * 8, hipMemcpyDeviceToHost));
  if (hs == LU_CAPACITY) d->owner = nullptr;  // the reach cache may be partial
  return lu_status_error(hs);
}

kh_status kh_sys_absproduct(kh_sys s, const double* x, double* y) {
  if (!s || !x || !y) return fail(KH_BAD_ARGUMENT, "null argument");
  sys_absproduct<double>(s, x, y);
  return KH_OK;
}

kh_status kh_sys_nnz(kh_sys s, int64_t* a, int64_t* l, int64_t* u) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  return sys_nnz<double>(s, a, l, u);
}

kh_status kh_sys_pivots(kh_sys s, int32_t* q, int32_t* prow) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  return sys_pivots<double>(s, q, prow);
}

void kh_sys_destroy(kh_sys s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  for (auto& d : g_scratch)
    if (d.owner == s) d.owner = nullptr;
  delete s;
}

// ---- complex seam: CSparse behind CMatrix (solver/matrix.go:39-46) -----------------------
kh_status kh_csys_create(int32_t n, kh_csys* out) {
  if (!out) return fail(KH_BAD_ARGUMENT, "null handle");
  if (n <= 0 || n > (1 << 22)) return fail(KH_BADSIZE, "size " + std::to_string(n));
  auto* s = new kh_csys_s();
  s->n = n;
  *out = s;
  return KH_OK;
}

kh_status kh_csys_clear(kh_csys s) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  std::fill(s->val.begin(), s->val.end(), kcplx{0.0, 0.0});
  s->valid = false;
  return KH_OK;
}

kh_status kh_csys_add_batch(kh_csys s, const int32_t* i, const int32_t* j, const double* v, int64_t count) {
  if (!s || (count > 0 && (!i || !j || !v))) return fail(KH_BAD_ARGUMENT, "null argument");
  return sys_add<kcplx>(s, i, j, reinterpret_cast<const kcplx*>(v), count);
}

kh_status kh_csys_solve(kh_csys s, const double* b, double* x) {
  if (!s || !b || !x) return fail(KH_BAD_ARGUMENT, "null argument");
  return sys_solve<kcplx>(s, reinterpret_cast<const kcplx*>(b), reinterpret_cast<kcplx*>(x));
}

kh_status kh_csys_absproduct(kh_csys s, const double* x, double* y) {
  if (!s || !x || !y) return fail(KH_BAD_ARGUMENT, "null argument");
  sys_absproduct<kcplx>(s, reinterpret_cast<const kcplx*>(x), y);
  return KH_OK;
}

kh_status kh_csys_nnz(kh_csys s, int64_t* a, int64_t* l, int64_t* u) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  return sys_nnz<kcplx>(s, a, l, u);
}

kh_status kh_csys_pivots(kh_csys s, int32_t* q, int32_t* prow) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  return sys_pivots<kcplx>(s, q, prow);
}

void kh_csys_destroy(kh_csys s) { delete s; }

// ---- batched numeric LU ------------------------------------------------------------------
kh_status kh_lu_batch_create(int32_t n, int64_t nnz, const int32_t* ri, const int32_t* cj, int32_t B,
                             kh_lu_batch* out) {
  if (!out || !ri || !cj || B <= 0 || nnz <= 0) return fail(KH_BAD_ARGUMENT, "bad argument");
  if (n <= 0 || n > (1 << 22)) return fail(KH_BADSIZE, "size " + std::to_string(n));
  std::vector<int32_t> r(ri, ri + nnz), c(cj, cj + nnz), slots(nnz);
  std::unordered_map<int64_t, int> seen;
  for (int64_t k = 0; k < nnz; k++) {
    if (r[k] < 0 || c[k] < 0 || r[k] >= n || c[k] >= n) return fail(KH_BAD_ARGUMENT, "entry out of range");
    if (!seen.emplace((int64_t)r[k] * n + c[k], 1).second) return fail(KH_BAD_ARGUMENT, "duplicate entry");
    slots[k] = (int32_t)k;
  }
  auto h = std::make_unique<kh_lu_batch_s>();
  h->n = n;
  h->nnz = nnz;
  h->B = B;
  h->pat = analyze_pattern(n, r, c, slots);
  h->pats.host.add(h->pat, slots_key(slots));
  KH_HIP(h->pats.sync());
  KH_HIP(h->work.ensure(n, B, lu_capacity(h->pat.fill_estimate, n)));
  KH_HIP(h->pid.alloc((size_t)B * 4));
  KH_HIP(h->status.alloc((size_t)B * 4));
  *out = h.release();
  return KH_OK;
}

kh_status kh_lu_batch_solve(kh_lu_batch h, const double* d_vals, const double* d_b, double* d_x, int32_t* d_status,
                            int32_t* d_prow, void* stream) {
  if (!h || !d_vals || !d_b || !d_x) return fail(KH_BAD_ARGUMENT, "null argument");
  LuArgs a = {};
  a.n = h->n;
  a.nslots = (int32_t)h->nnz;
  a.B = h->B;
  a.pats = h->pats.desc.as<PatDesc>();
  a.pool = h->pats.pool.as<int32_t>();
  a.pid = h->pid.as<int32_t>();
  a.val = d_vals;
  a.rhs = d_b;
  a.x = d_x;
  h->work.fill(a);
  hipStream_t st = static_cast<hipStream_t>(stream);
  int32_t* status = d_status ? d_status : h->status.as<int32_t>();
  hipLaunchKernelGGL(k_lu, dim3(grid_for(h->B)), dim3(kBlock), 0, st, a, nullptr, nullptr, status);
  KH_HIP(hipGetLastError());
  if (d_prow)
    KH_HIP(hipMemcpyAsync(d_prow, h->work.prow.p, (size_t)h->n * h->B * 4, hipMemcpyDeviceToDevice, st));
  return KH_OK;
}

kh_status kh_lu_batch_solve_mixed(kh_lu_batch h, const double* d_vals, const double* d_b, double* d_x,
                                  int32_t* d_status, int32_t* d_prow, int32_t* d_passes, void* stream) {
  if (!h || !d_vals || !d_b || !d_x) return fail(KH_BAD_ARGUMENT, "null argument");
  LuArgs a = {};
  a.n = h->n;
  a.nslots = (int32_t)h->nnz;
  a.B = h->B;
  a.pats = h->pats.desc.as<PatDesc>();
  a.pool = h->pats.pool.as<int32_t>();
  a.pid = h->pid.as<int32_t>();
  a.val = d_vals;
  a.rhs = d_b;
  a.x = d_x;
  h->work.fill(a);
  hipStream_t st = static_cast<hipStream_t>(stream);
  int32_t* status = d_status ? d_status : h->status.as<int32_t>();
  hipLaunchKernelGGL(k_lu_mixed, dim3(grid_for(h->B)), dim3(kBlock), 0, st, a, status, d_passes);
  KH_HIP(hipGetLastError());
  if (d_prow)
    KH_HIP(hipMemcpyAsync(d_prow, h->work.prow.p, (size_t)h->n * h->B * 4, hipMemcpyDeviceToDevice, st));
  return KH_OK;
}

kh_status kh_lu_batch_info(kh_lu_batch h, int32_t* q, int64_t* nnzL, int64_t* nnzU, int64_t* capL, int64_t* capU) {
  if (!h) return fail(KH_BAD_ARGUMENT, "null handle");
  if (q) std::memcpy(q, h->pat.q.data(), (size_t)h->n * 4);
  if (nnzL || nnzU) {
    int32_t lpn = 0, upn = 0;
    KH_HIP(hipMemcpy(&lpn, h->work.lp.as<int32_t>() + (size_t)h->n * h->B, 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(&upn, h->work.up.as<int32_t>() + (size_t)h->n * h->B, 4, hipMemcpyDeviceToHost));
    if (nnzL) *nnzL = lpn;
    if (nnzU) *nnzU = (int64_t)upn + h->n;
  }
  if (capL) *capL = h->work.capL;
  if (capU) *capU = h->work.capU;
  return KH_OK;
}

void kh_lu_batch_destroy(kh_lu_batch h) { delete h; }

// ---- topologies and batched analyses -------------------------------------------------------
kh_status kh_topology_create(int32_t num_nodes, const kh_element* elems, int32_t n_elems, int32_t nparams,
                             const double* params_c0, kh_topology* out) {
  if (!out || !elems || n_elems <= 0 || !params_c0 || num_nodes < 1) return fail(KH_BAD_ARGUMENT, "bad argument");
  auto t = std::make_unique<kh_topology_s>();
  if (!t->t.build(num_nodes, elems, n_elems, nparams, params_c0)) return fail(KH_UNSUPPORTED, t->t.error);
  const Topology& T = t->t;
  KH_HIP(t->et.alloc(T.et.size() * 4));
  KH_HIP(hipMemcpy(t->et.p, T.et.data(), T.et.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->isV.alloc(T.isV.size() * 4));
  KH_HIP(hipMemcpy(t->isV.p, T.isV.data(), T.isV.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->slot_i.alloc(T.slot_i.size() * 4));
  KH_HIP(hipMemcpy(t->slot_i.p, T.slot_i.data(), T.slot_i.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->slot_j.alloc(T.slot_j.size() * 4));
  KH_HIP(hipMemcpy(t->slot_j.p, T.slot_j.data(), T.slot_j.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->diag.alloc(T.diag_slot.size() * 4));
  KH_HIP(hipMemcpy(t->diag.p, T.diag_slot.data(), T.diag_slot.size() * 4, hipMemcpyHostToDevice));
  std::vector<int32_t> rm(T.nslots);  // the dense backend's AbsProduct order (row by row)
  for (int32_t k = 0; k < T.nslots; k++) rm[k] = k;
  std::sort(rm.begin(), rm.end(), [&](int32_t a, int32_t b) {
    return T.slot_i[a] != T.slot_i[b] ? T.slot_i[a] < T.slot_i[b] : T.slot_j[a] < T.slot_j[b];
  });
  KH_HIP(t->rowmajor.alloc(rm.size() * 4));
  KH_HIP(hipMemcpy(t->rowmajor.p, rm.data(), rm.size() * 4, hipMemcpyHostToDevice));
  std::vector<int32_t> li(T.nl_idx);
  li.insert(li.end(), T.rx_idx.begin(), T.rx_idx.end());
  li.insert(li.end(), T.src_idx.begin(), T.src_idx.end());
  KH_HIP(t->lists.alloc(std::max<size_t>(li.size(), 1) * 4));
  if (!li.empty()) KH_HIP(hipMemcpy(t->lists.p, li.data(), li.size() * 4, hipMemcpyHostToDevice));
  {
    const WavePlan w = build_wave_plan(T, kWaveRecordLimit);
    const std::vector<int32_t>* parts[10] = {&w.ch_el,        &w.ch_t,          &w.tgt,          &w.tel_ptr,
                                             &w.tel,          &w.rec_off,       &w.ro_ptr_sparse, &w.ro_slot_sparse,
                                             &w.ro_ptr_dense, &w.ro_slot_dense};
    std::vector<int32_t> all;
    for (int k = 0; k < 10; k++) {
      t->wave_off[k] = (int64_t)all.size();
      all.insert(all.end(), parts[k]->begin(), parts[k]->end());
    }
    t->nchunks = (int32_t)w.ch_el.size() - 1;
    t->rec_cap = w.rec_cap;
    KH_HIP(t->wave.alloc(std::max<size_t>(all.size(), 1) * 4));
    if (!all.empty()) KH_HIP(hipMemcpy(t->wave.p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
  }
  *out = t.release();
  return KH_OK;
}

int32_t kh_topology_size(kh_topology t) { return t ? t->t.n : -1; }

void kh_topology_destroy(kh_topology t) { delete t; }

}  // extern "C"

// ---- the batch driver -----------------------------------------------------------------------
// The compiled k_persist variants (kh_persist.hip, one translation unit per
// register budget and kind set).
namespace kh {
#define KH_PERSIST_VARIANT(W, K)                                                                          \
  hipError_t kh_launch_persist_w##W##_k##K(unsigned, size_t, const Group*, const int32_t*, int32_t, int32_t*);
#define KH_PERSIST_KINDS(W) KH_PERSIST_VARIANT(W, 126) KH_PERSIST_VARIANT(W, 190) KH_PERSIST_VARIANT(W, 318) \
  KH_PERSIST_VARIANT(W, 32766)
KH_PERSIST_KINDS(1)
KH_PERSIST_KINDS(2)
KH_PERSIST_KINDS(4)
#undef KH_PERSIST_KINDS
#undef KH_PERSIST_VARIANT

#define KH_WAVE_VARIANT(K)                                                                                 \
  hipError_t kh_launch_wave_k##K(unsigned, size_t, const Group*, const int32_t*, const int32_t*, int32_t, int32_t*, \
                                 int32_t, int32_t*);
KH_WAVE_VARIANT(126)
KH_WAVE_VARIANT(190)
KH_WAVE_VARIANT(318)
KH_WAVE_VARIANT(32766)
#undef KH_WAVE_VARIANT

hipError_t kh_launch_wave(unsigned kinds, unsigned grid, size_t lds_bytes, const Group* G, const int32_t* gw,
                          const int32_t* cw, int32_t nwork, int32_t* next, int32_t max_solves, int32_t* counters) {
  using L = hipError_t (*)(unsigned, size_t, const Group*, const int32_t*, const int32_t*, int32_t, int32_t*, int32_t,
                           int32_t*);
  static const L table[4] = {kh_launch_wave_k126, kh_launch_wave_k190, kh_launch_wave_k318, kh_launch_wave_k32766};
  const unsigned v = persist_kinds(kinds);
  const int kv = v == kKindsD ? 0 : v == kKindsM ? 1 : v == kKindsQ ? 2 : 3;
  return table[kv](grid, lds_bytes, G, gw, cw, nwork, next, max_solves, counters);
}

hipError_t kh_launch_persist(int waves, unsigned kinds, unsigned grid, size_t lds_bytes, const Group* G,
                             const int32_t* bgroup, int32_t max_solves, int32_t* counters) {
  using L = hipError_t (*)(unsigned, size_t, const Group*, const int32_t*, int32_t, int32_t*);
  static const L table[3][4] = {
      {kh_launch_persist_w1_k126, kh_launch_persist_w1_k190, kh_launch_persist_w1_k318, kh_launch_persist_w1_k32766},
      {kh_launch_persist_w2_k126, kh_launch_persist_w2_k190, kh_launch_persist_w2_k318, kh_launch_persist_w2_k32766},
      {kh_launch_persist_w4_k126, kh_launch_persist_w4_k190, kh_launch_persist_w4_k318, kh_launch_persist_w4_k32766}};
  const unsigned v = persist_kinds(kinds);
  const int kv = v == kKindsD ? 0 : v == kKindsM ? 1 : v == kKindsQ ? 2 : 3;
  const int wv = waves == 4 ? 2 : waves == 2 ? 1 : 0;
  return table[wv][kv](grid, lds_bytes, G, bgroup, max_solves, counters);
}
}  // namespace kh


namespace {

// Device state of the circuits of one topology, stored per circuit.
struct GroupRun {
  kh_topology_s* topo = nullptr;
  int64_t B = 0;
  int n = 0;
  int32_t gid = 0;
  DBuf par, lane, nodeset, st, val, present, rhs, x, xold, xsave, ptprev, xacc, absbuf, ci, cd, lustatus, pid, hash, solves;
  DBuf swx, swit;           // DC sweep: per circuit [points][n] solutions, [points] Newton iterations
  DBuf swi;                 // DC sweep (when asked for): per circuit [points][nel] element currents
  bool want_sw_i = false;
  int64_t sw_points = 0;
  DevPatterns pats;
  LuWork work;
  bool has_lane = false;
  bool has_nodeset = false;
  bool dense = false;  // Options.Backend resolved for this topology's size
  kh_stats stats = {};

  kh_status setup(const double* d_params, const double* d_lane_spec, const double* d_nodeset) {
    const Topology& T = topo->t;
    n = T.n;
    if (dense) KH_HIP(work.ensure_dense(n, B));
    size_t nd = (size_t)n * B * 8;
    KH_HIP(par.alloc((size_t)T.nparams * B * 8));
    KH_HIP(st.alloc((size_t)std::max(T.nstate, 1) * B * 8));
    KH_HIP(val.alloc((size_t)T.nslots * B * 8));
    KH_HIP(present.alloc((size_t)T.nslots * B));
    KH_HIP(rhs.alloc(nd)); KH_HIP(x.alloc(nd)); KH_HIP(xold.alloc(nd)); KH_HIP(xsave.alloc(nd));
    KH_HIP(ptprev.alloc(nd)); KH_HIP(xacc.alloc(nd)); KH_HIP(absbuf.alloc(nd));
    KH_HIP(ci.alloc((size_t)NCI * B * 4));
    KH_HIP(cd.alloc((size_t)NCD * B * 8));
    KH_HIP(lustatus.alloc((size_t)B * 4));
    KH_HIP(pid.alloc((size_t)B * 4));
    KH_HIP(hash.alloc((size_t)B * 8));
    KH_HIP(solves.alloc((size_t)B * 4));
    int64_t tot = (int64_t)T.nparams * B;
    hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, d_params,
                       par.as<double>(), (int64_t)T.nparams, B);
    KH_HIP(hipGetLastError());
    has_lane = d_lane_spec != nullptr;
    if (has_lane) {
      KH_HIP(lane.alloc((size_t)4 * B * 8));
      hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((4 * B + 255) / 256)), dim3(256), 0, 0, d_lane_spec,
                         lane.as<double>(), (int64_t)4, B);
      KH_HIP(hipGetLastError());
    }
    if (sw_points > 0) {
      KH_HIP(swx.alloc((size_t)B * sw_points * T.n * 8));
      KH_HIP(swit.alloc((size_t)B * sw_points * 4));
      if (want_sw_i) KH_HIP(swi.alloc((size_t)B * sw_points * T.nel * 8));
    }
    has_nodeset = d_nodeset != nullptr && T.num_nodes > 1;
    if (has_nodeset) {
      int64_t rows = T.num_nodes - 1;
      KH_HIP(nodeset.alloc((size_t)rows * B * 8));
      hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((rows * B + 255) / 256)), dim3(256), 0, 0, d_nodeset,
                         nodeset.as<double>(), rows, B);
      KH_HIP(hipGetLastError());
    }
    return KH_OK;
  }

  CtlArgs ctl_args(const CtlParams& p) const {
    const Topology& T = topo->t;
    CtlArgs a = {};
    a.ckt.n = T.n; a.ckt.nel = T.nel; a.ckt.nslots = T.nslots; a.ckt.numNodes = T.num_nodes;
    a.ckt.et = topo->et.as<int32_t>(); a.ckt.isV = topo->isV.as<int32_t>();
    a.ckt.slot_i = topo->slot_i.as<int32_t>(); a.ckt.slot_j = topo->slot_j.as<int32_t>();
    a.ckt.diag_slot = topo->diag.as<int32_t>();
    a.ckt.nonlin = T.nonlin ? 1 : 0;
    a.ckt.dense = dense ? 1 : 0;
    a.ckt.order = dense ? topo->rowmajor.as<int32_t>() : nullptr;
    a.ckt.n_nl = (int32_t)T.nl_idx.size();
    a.ckt.n_rx = (int32_t)T.rx_idx.size();
    a.ckt.n_src = (int32_t)T.src_idx.size();
    {
      const int32_t* w = topo->wave.as<int32_t>();
      const int64_t* o = topo->wave_off;
      a.ckt.nchunks = topo->nchunks;
      a.ckt.rec_cap = topo->rec_cap;
      a.ckt.ch_el = w + o[0]; a.ckt.ch_t = w + o[1]; a.ckt.tgt = w + o[2]; a.ckt.tel_ptr = w + o[3];
      a.ckt.tel = w + o[4]; a.ckt.rec_off = w + o[5];
      a.ckt.ro_ptr = w + (dense ? o[8] : o[6]);
      a.ckt.ro_slot = w + (dense ? o[9] : o[7]);
    }
    a.ckt.nl = topo->lists.as<int32_t>();
    a.ckt.rx = a.ckt.nl + a.ckt.n_nl;
    a.ckt.src = a.ckt.rx + a.ckt.n_rx;
    a.p = p;
    a.B = B;
    a.par = par.as<double>();
    a.st = st.as<double>(); a.val = val.as<double>(); a.present = present.as<uint8_t>();
    a.rhs = rhs.as<double>(); a.x = x.as<double>(); a.xold = xold.as<double>();
    a.xsave = xsave.as<double>(); a.ptprev = ptprev.as<double>(); a.xacc = xacc.as<double>();
    a.absbuf = absbuf.as<double>();
    a.ci = ci.as<int32_t>(); a.cd = cd.as<double>();
    a.lu_status = lustatus.as<int32_t>();
    a.counters = nullptr;
    a.lane_spec = has_lane ? lane.as<double>() : nullptr;
    a.nodeset = has_nodeset ? nodeset.as<double>() : nullptr;
    a.sw_x = sw_points ? swx.as<double>() : nullptr;
    a.sw_it = sw_points ? swit.as<int32_t>() : nullptr;
    a.sw_i = (sw_points && want_sw_i) ? swi.as<double>() : nullptr;
    return a;
  }

  LuArgs lu_args() const {
    const Topology& T = topo->t;
    LuArgs a = {};
    a.n = T.n;
    a.nslots = T.nslots;
    a.B = B;
    a.pats = pats.desc.as<PatDesc>();
    a.pool = pats.pool.as<int32_t>();
    a.pid = pid.as<int32_t>();
    a.val = val.as<double>();
    a.rhs = rhs.as<double>();
    a.x = x.as<double>();
    work.fill(a);
    return a;
  }

  // Symbolic phase for circuits whose stamps grew their pattern: one ordering
  // per distinct pattern, shared by every circuit that has it.
  kh_status new_patterns() {
    const Topology& T = topo->t;
    hipLaunchKernelGGL(k_hash, dim3(grid_for(B)), dim3(kBlock), 0, 0, present.as<uint8_t>(), T.nslots, B,
                       ci.as<int32_t>(), hash.as<uint64_t>());
    KH_HIP(hipGetLastError());
    std::vector<int32_t> wait(B), hpid(B);
    std::vector<uint64_t> hh(B);
    // The control words come back whole (a strided copy of one word per
    // circuit costs milliseconds); the wait words are cleared in that copy
    // and it goes back whole.
    std::vector<int32_t> hci((size_t)NCI * B);
    KH_HIP(hipMemcpy(hci.data(), ci.p, hci.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < B; c++) wait[c] = hci[(size_t)c * NCI + CI_WAIT];
    KH_HIP(hipMemcpy(hh.data(), hash.p, (size_t)B * 8, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hpid.data(), pid.p, (size_t)B * 4, hipMemcpyDeviceToHost));
    std::unordered_map<uint64_t, int32_t> by_hash;
    std::vector<uint8_t> col(T.nslots);
    bool grow = false;
    for (int64_t c = 0; c < B; c++) {
      if (!wait[c]) continue;
      if (wait[c] == WAIT_CAPACITY) grow = true;  // same pattern, needs more L/U room
      auto it = by_hash.find(hh[c]);
      if (it != by_hash.end()) {
        hpid[c] = it->second;
        continue;
      }
      KH_HIP(hipMemcpy(col.data(), present.as<uint8_t>() + c * T.nslots, T.nslots, hipMemcpyDeviceToHost));
      std::vector<int32_t> slots;
      for (int s = 0; s < T.nslots; s++)
        if (col[s]) slots.push_back(s);
      std::string key = slots_key(slots);
      int32_t id;
      auto f = pats.host.by_key.find(key);
      if (f != pats.host.by_key.end()) {
        id = f->second;
      } else {
        Pattern P = analyze_pattern(T.n, T.slot_i, T.slot_j, slots);
        id = pats.host.add(P, key);
        stats.analyses++;
      }
      by_hash[hh[c]] = id;
      hpid[c] = id;
    }
    for (int64_t c = 0; c < B; c++) hci[(size_t)c * NCI + CI_WAIT] = 0;
    stats.patterns = (int64_t)pats.host.desc.size();
    KH_HIP(pats.sync());
    int64_t cap = std::max(work.capL, lu_capacity_batch(pats.host.max_fill, T.n));
    if (grow) cap = std::max(cap, 2 * work.capL);
    KH_HIP(work.ensure(T.n, B, cap));
    KH_HIP(hipMemcpy(pid.p, hpid.data(), (size_t)B * 4, hipMemcpyHostToDevice));
    KH_HIP(hipMemcpy(ci.p, hci.data(), hci.size() * 4, hipMemcpyHostToDevice));
    return KH_OK;
  }

  // Element reports rep [B][nel][RP_FIELDS] and ledgers led [B][LG_FIELDS]
  // (both optional; rep is computed into scratch when only led is asked for)
  // at each circuit's final solution and state.
  kh_status report(const CtlParams& p, double* d_rep, double* d_led) {
    if (!d_rep && !d_led) return KH_OK;
    const Topology& T = topo->t;
    Group g = {};
    g.ctl = ctl_args(p);
    g.lu = lu_args();
    g.npar = T.nparams;
    g.nstate = std::max(T.nstate, 1);
    g.cpb = 1;
    DBuf tmp;
    double* rep = d_rep;
    if (!rep) {
      KH_HIP(tmp.alloc((size_t)B * T.nel * RP_FIELDS * 8));
      rep = tmp.as<double>();
    }
    const int64_t tot = B * (int64_t)T.nel;
    if (tot > 0) {
      hipLaunchKernelGGL(k_report, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, g, rep);
      KH_HIP(hipGetLastError());
    }
    if (d_led) {
      hipLaunchKernelGGL(k_report_ledger, dim3((unsigned)((B + 63) / 64)), dim3(64), 0, 0, g, rep, d_led);
      KH_HIP(hipGetLastError());
    }
    KH_HIP(hipDeviceSynchronize());
    return KH_OK;
  }

  // Per-circuit results (caller's SoA layout) and batch counters.
  kh_status outputs(double* d_x, int32_t* d_status, int32_t* d_iters, int32_t* d_acc, int32_t* d_rej,
                    double* d_ledger, double* d_sw_x = nullptr, int32_t* d_sw_it = nullptr,
                    int32_t* d_points = nullptr, double* d_sw_i = nullptr) {
    if (sw_points > 0 && d_sw_i && want_sw_i) {  // [B][points*nel] -> [points*nel][B]
      const int64_t cols = (int64_t)sw_points * topo->t.nel;
      const int64_t tot = cols * B;
      hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, swi.as<double>(),
                         d_sw_i, B, cols);
      KH_HIP(hipGetLastError());
    }
    if (sw_points > 0 && d_sw_x) {  // [B][points*n] -> [points*n][B]
      int64_t tot = (int64_t)sw_points * n * B;
      hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, swx.as<double>(),
                         d_sw_x, B, (int64_t)sw_points * n);
      KH_HIP(hipGetLastError());
    }
    if (sw_points > 0 && d_sw_it) {
      int64_t tot = (int64_t)sw_points * B;
      hipLaunchKernelGGL(k_transpose<int32_t>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0,
                         swit.as<int32_t>(), d_sw_it, B, (int64_t)sw_points);
      KH_HIP(hipGetLastError());
    }
    if (d_x) {
      int64_t tot = (int64_t)n * B;
      hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, x.as<double>(),
                         d_x, B, (int64_t)n);
      KH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_outputs, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, 0, ci.as<int32_t>(),
                       cd.as<double>(), B, d_status, d_iters, d_acc, d_rej, d_ledger, solves.as<int32_t>(), d_points);
    KH_HIP(hipGetLastError());
    std::vector<int32_t> w(NCI * B);
    KH_HIP(hipMemcpy(w.data(), ci.p, (size_t)NCI * B * 4, hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < B; c++) {
      const int32_t* v = &w[c * NCI];
      stats.newton_iterations += v[CI_ITERS];
      stats.accepted_steps += v[CI_ACCEPTED];
      stats.rejected_steps += v[CI_REJECTED];
      stats.solves += v[CI_SOLVES];
    }
    std::vector<double> dv(NCD * B);
    KH_HIP(hipMemcpy(dv.data(), cd.p, (size_t)NCD * B * 8, hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < B; c++) {
      stats.cycles_lu += dv[c * NCD + CD_CYC_LU];
      stats.cycles_control += dv[c * NCD + CD_CYC_CTL];
      stats.refine_passes += (int64_t)dv[c * NCD + CD_REFINE];
      stats.fp64_redos += (int64_t)dv[c * NCD + CD_REDO];
      stats.cycles_stamp += dv[c * NCD + CD_CYC_STAMP];
      stats.cycles_converge += dv[c * NCD + CD_CYC_CONV];
      stats.cycles_lte += dv[c * NCD + CD_CYC_LTE];
      stats.cycles_accept += dv[c * NCD + CD_CYC_ACCEPT];
      stats.cycles_ledger += dv[c * NCD + CD_CYC_LEDGER];
    }
    // Algorithmic bytes per executed solve: matrix values and right-hand side
    // read once, solution written once, 8*(nnzA + 2n), plus the device
    // parameters read by the stamp, 8*P (largest pattern of the group:
    // patterns only grow).
    int64_t nnz = 0;
    for (const auto& d : pats.host.desc) nnz = std::max<int64_t>(nnz, d.nnz);
    // Mixed precision (FP32 factors): 4*(nnzA + 2n) per certified solve plus
    // 8*nnzA + 16n (FP64 residual) per further refinement pass; a system
    // redone in FP64 counts as an FP64 solve.
    const int64_t P = topo->t.nparams, n2 = 2 * (int64_t)n;
    const int64_t mixed = stats.solves - stats.fp64_redos;
    if (stats.refine_passes > 0)
      stats.bytes_algorithmic = mixed * (4 * (nnz + n2) + 8 * P) + (stats.refine_passes - mixed) * (8 * nnz + 8 * n2) +
                                stats.fp64_redos * 8 * (nnz + n2 + P);
    else
      stats.bytes_algorithmic = stats.solves * 8 * (nnz + n2 + P);
    return KH_OK;
  }
};

// Runs every group to completion with persistent launches.
struct Runner {
  std::vector<GroupRun*> groups;
  DBuf dgroups, dbgroup, counters;
  int32_t* hcount = nullptr;
  kh_stats stats = {};
  // Solves per circuit per launch. Launches are kept short (kLaunchTargetMs) so
  // that no single dispatch occupies the queue for long: the budget starts
  // small and adapts to the measured launch duration.
  int32_t max_solves = 4;
  static constexpr float kLaunchTargetMs = 25.0f;
  int32_t cpb = 1;
  // LDS staging (k_persist): each circuit stages what fits its share of the
  // LDS (plan_staging): the share keeps every workgroup of the launch resident
  // at once (a compute unit's LDS split over the workgroups it must hold,
  // at most lds_limit per workgroup).
  bool lds_staging = false;
  int32_t waves = 1;  // k_persist variant: resident wavefronts per SIMD (1, 2 or 4)
  int64_t lds_limit = 64 * 1024;  // raised to the device's opt-in maximum (160 KiB on gfx950)
  int64_t lds_per_cu = 0;         // LDS of one compute unit (0: unknown, lds_limit assumed)
  int32_t ncu = 0;                // compute units (0: unknown)
  size_t lds_launch = 0;
  // k_wave (kh_options.circuit_wave): one wavefront per circuit, circuits
  // taken from a work list, largest first.
  bool wave = false;
  DBuf dgw, dcw, dnext;

  ~Runner() {
    if (hcount) (void)hipHostFree(hcount);
  }

  kh_status run(const CtlParams& p) {
    size_t ng = groups.size();
    unsigned kinds = 0;  // element kinds of the whole batch: picks the k_persist variant
    for (auto* gr : groups) kinds |= gr->topo->t.kinds;
    std::vector<int32_t> bg;
    std::vector<Group> hg(ng);
    unsigned wave_grid = 0;
    int32_t nwork = 0;
    if (wave) {
      // Work list: every circuit, the largest topologies first (the longest
      // serial chains start first; the small ones fill in behind them).
      std::vector<std::pair<int32_t, int32_t>> order;
      for (size_t g = 0; g < ng; g++) order.push_back({groups[g]->topo->t.n, (int32_t)g});
      std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
      std::vector<int32_t> gw, cw;
      for (const auto& o : order)
        for (int64_t c = 0; c < groups[o.second]->B; c++) {
          gw.push_back(o.second);
          cw.push_back((int32_t)c);
        }
      nwork = (int32_t)gw.size();
      KH_HIP(dgw.alloc(gw.size() * 4));
      KH_HIP(dcw.alloc(cw.size() * 4));
      KH_HIP(dnext.alloc(8));
      KH_HIP(hipMemcpy(dgw.p, gw.data(), gw.size() * 4, hipMemcpyHostToDevice));
      KH_HIP(hipMemcpy(dcw.p, cw.data(), cw.size() * 4, hipMemcpyHostToDevice));
      int64_t lds = 0;
      for (size_t g = 0; g < ng; g++)
        lds = std::max(lds, wave_lds_bytes(groups[g]->topo->t.n, groups[g]->topo->rec_cap));
      lds = round8(lds);
      int dev = 0, optin = 0, ncu = 0, per_cu = 0;
      KH_HIP(hipGetDevice(&dev));
      KH_HIP(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
      KH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
      KH_HIP(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
      if (lds > optin)
        return fail(KH_CAPACITY, "circuit_wave: a circuit of this size needs " + std::to_string(lds) +
                                     " bytes of LDS per wavefront, above the device's " + std::to_string(optin));
      lds_launch = (size_t)lds;
      const int64_t resident = (int64_t)std::max(ncu, 1) *
                               std::min<int64_t>(16, std::max<int64_t>(1, (per_cu > 0 ? per_cu : optin) / lds));
      wave_grid = (unsigned)std::min<int64_t>(nwork, resident);
    }
    for (size_t g = 0; g < ng; g++) {
      hg[g].block0 = (int64_t)bg.size();
      hg[g].gid = (int32_t)g;
      hg[g].cpb = cpb;
      hg[g].npar = groups[g]->topo->t.nparams;
      hg[g].nstate = std::max(groups[g]->topo->t.nstate, 1);
      int64_t nb = (groups[g]->B + cpb - 1) / cpb;
      for (int64_t k = 0; k < nb; k++) bg.push_back((int32_t)g);
    }
    KH_HIP(dbgroup.alloc(bg.size() * 4));
    KH_HIP(hipMemcpy(dbgroup.p, bg.data(), bg.size() * 4, hipMemcpyHostToDevice));
    KH_HIP(dgroups.alloc(ng * sizeof(Group)));
    KH_HIP(counters.alloc(ng * 8));
    KH_HIP(hipHostMalloc((void**)&hcount, ng * 8, hipHostMallocDefault));
    hipEvent_t ev[2];
    for (auto& e : ev) KH_HIP(hipEventCreate(&e));
    struct EvGuard {
      hipEvent_t* e;
      ~EvGuard() { for (int k = 0; k < 2; k++) (void)hipEventDestroy(e[k]); }
    } guard{ev};
    auto t0 = std::chrono::steady_clock::now();
    // KH_PROGRESS=<seconds>: a line on stderr at that interval while the
    // batch runs (long batches under a watchdog).
    const char* pe = std::getenv("KH_PROGRESS");
    const double progress_s = pe ? std::atof(pe) : 0.0;
    double next_report = progress_s;
    bool dirty = true;
    for (;;) {
      if (dirty) {  // group descriptors: LU workspace pointers change when capacity grows
        if (!wave) lds_launch = 0;
        // workgroups each compute unit holds when the whole launch is resident
        int64_t wg_cu = ncu > 0 ? ((int64_t)bg.size() + ncu - 1) / ncu : 1;
        wg_cu = std::min<int64_t>(std::max<int64_t>(wg_cu, 1), 4 * (int64_t)waves);
        const int64_t wg_lds = std::min<int64_t>(lds_limit, (lds_per_cu > 0 ? lds_per_cu : lds_limit) / wg_cu);
        for (size_t g = 0; g < ng; g++) {
          hg[g].ctl = groups[g]->ctl_args(p);
          hg[g].lu = groups[g]->lu_args();
          hg[g].lds_bytes = 0;
          hg[g].stage_all = 0;
          hg[g].stage_mask = 0;
          if (lds_staging && !wave) {
            plan_staging(hg[g], wg_lds / hg[g].cpb);
            if (hg[g].lds_bytes > 0)
              lds_launch = std::max(lds_launch, (size_t)hg[g].lds_bytes * (size_t)hg[g].cpb);
          }
        }
        stats.staged_bytes = 0;
        stats.staged_partial = 0;
        for (size_t g = 0; g < ng; g++) {
          stats.staged_bytes = std::max<int64_t>(stats.staged_bytes, hg[g].lds_bytes);
          int nar = 0;
          for_each_array(hg[g], [&](const auto&, int64_t, bool, bool, int) { nar++; });
          const uint64_t full = nar >= 64 ? ~0ull : ((1ull << nar) - 1);
          if (hg[g].lds_bytes > 0 && hg[g].stage_mask != full) stats.staged_partial++;
        }
        KH_HIP(hipMemcpy(dgroups.p, hg.data(), ng * sizeof(Group), hipMemcpyHostToDevice));
        dirty = false;
      }
      KH_HIP(hipMemsetAsync(counters.p, 0, ng * 8, 0));
      if (wave) KH_HIP(hipMemsetAsync(dnext.p, 0, 8, 0));
      KH_HIP(hipEventRecord(ev[0], 0));
      if (wave)
        KH_HIP(kh_launch_wave(kinds, wave_grid, lds_launch, dgroups.as<Group>(), dgw.as<int32_t>(), dcw.as<int32_t>(),
                              nwork, dnext.as<int32_t>(), max_solves, counters.as<int32_t>()));
      else
        KH_HIP(kh_launch_persist(waves, kinds, (unsigned)bg.size(), lds_launch, dgroups.as<Group>(),
                                 dbgroup.as<int32_t>(), max_solves, counters.as<int32_t>()));
      KH_HIP(hipEventRecord(ev[1], 0));
      KH_HIP(hipMemcpyAsync(hcount, counters.p, ng * 8, hipMemcpyDeviceToHost, 0));
      KH_HIP(hipStreamSynchronize(0));
      float ms = 0;
      KH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      stats.lu_seconds += ms * 1e-3;
      stats.lu_launches++;
      stats.ticks++;
      if (ms < 0.5f * kLaunchTargetMs && max_solves < (1 << 14)) max_solves *= 2;
      else if (ms > 2.0f * kLaunchTargetMs && max_solves > 1) max_solves /= 2;
      int64_t running = 0;
      for (size_t g = 0; g < ng; g++) {
        running += hcount[2 * g];
        if (hcount[2 * g + 1] > 0) {
          kh_status s = groups[g]->new_patterns();
          if (s != KH_OK) return s;
          dirty = true;
        }
      }
      if (running == 0) break;
      if (progress_s > 0) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el >= next_report) {
          fprintf(stderr, "kh_batch_run: %.0f s, %lld launches, %lld circuits running\n", el,
                  (long long)stats.ticks, (long long)running);
          next_report = el + progress_s;
        }
      }
    }
    KH_HIP(hipDeviceSynchronize());
    stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return KH_OK;
  }

  void total() {
    for (GroupRun* g : groups) {
      stats.solves += g->stats.solves;
      stats.newton_iterations += g->stats.newton_iterations;
      stats.accepted_steps += g->stats.accepted_steps;
      stats.rejected_steps += g->stats.rejected_steps;
      stats.analyses += g->stats.analyses;
      stats.patterns += g->stats.patterns;
      stats.bytes_algorithmic += g->stats.bytes_algorithmic;
      stats.cycles_lu += g->stats.cycles_lu;
      stats.cycles_control += g->stats.cycles_control;
      stats.refine_passes += g->stats.refine_passes;
      stats.fp64_redos += g->stats.fp64_redos;
      stats.cycles_stamp += g->stats.cycles_stamp;
      stats.cycles_converge += g->stats.cycles_converge;
      stats.cycles_lte += g->stats.cycles_lte;
      stats.cycles_accept += g->stats.cycles_accept;
      stats.cycles_ledger += g->stats.cycles_ledger;
    }
  }
};

CtlParams resolve(const kh_options* o, int analysis) {
  CtlParams p = {};
  p.relTol = (o && o->reltol > 0) ? o->reltol : 1e-3;
  p.absTol = (o && o->abstol > 0) ? o->abstol : 1e-12;
  p.vnTol = (o && o->vntol > 0) ? o->vntol : 1e-6;
  p.gmin = (o && o->gmin > 0) ? o->gmin : 1e-12;
  p.ledgerTol = (o && o->ledger_tol > 0) ? o->ledger_tol : 1e-9;
  p.itl1 = (o && o->itl1 > 0) ? o->itl1 : 100;
  p.itl4 = (o && o->itl4 > 0) ? o->itl4 : 10;
  p.noGmin = o ? o->no_gmin_stepping : 0;
  p.noSource = o ? o->no_source_stepping : 0;
  p.noPT = o ? o->no_pseudo_transient : 0;
  p.skipEnergy = o ? o->skip_energy_check : 0;
  p.analysis = analysis;
  p.precision = o ? o->precision : 0;
  return p;
}

kh_status tran_params(const kh_tran_spec* spec, CtlParams& p) {
  // Tran's spec validation, analysis/tran.go:71-74
  if (!(spec->step > 0) || !(spec->stop > 0) || spec->start < 0 || spec->start >= spec->stop || spec->max_step < 0 ||
      std::isnan(spec->step + spec->stop + spec->start + spec->max_step) || std::isinf(spec->stop))
    return fail(KH_BAD_TRAN_SPEC, "invalid transient specification");
  p.step = spec->step;
  p.stop = spec->stop;
  p.start = spec->start;
  p.maxStep = spec->max_step > 0 ? spec->max_step : std::min(spec->step, (spec->stop - spec->start) / 50);
  p.chgTol = spec->chgtol > 0 ? spec->chgtol : 1e-14;
  p.trTol = spec->trtol > 0 ? spec->trtol : 7;
  p.minStep = spec->min_step > 0 ? spec->min_step : 1e-12 * spec->stop;
  p.maxSteps = spec->max_steps > 0 ? spec->max_steps : 1000000;
  p.uic = spec->uic;
  p.method = spec->method == 1 ? BE : TRAP;
  return KH_OK;
}

std::mutex g_run_mu;  // batch entry points run one at a time per process

// Outputs of a DC sweep (one group).
struct SweepOut {
  double* d_x = nullptr;
  double* d_i = nullptr;  // optional element currents per point
  int32_t* d_iters = nullptr;
  int32_t* d_points = nullptr;
};

// Runs every group of a batch to completion; p is the resolved analysis.
// The group runs are handed to `keep` (when given) for a further phase on
// their per-circuit state; the caller then holds g_run_mu itself.
// Backend.Resolve (solver/matrix.go:77-85): Auto is dense below 30 unknowns.
constexpr int kDenseThreshold = 30, kMaxDense = 4096;  // DenseThreshold, MaxDense (solver/dense.go:12)
bool use_dense(const kh_options* opt, int n) {
  const int b = opt ? opt->backend : 0;
  return b == 1 || (b == 0 && n < kDenseThreshold);
}

kh_status batch_run(int32_t ngroups, const kh_group* groups, const kh_options* opt, const CtlParams& p,
                    const kh_tran_spec* spec, const SweepOut* sw, kh_stats* stats,
                    std::vector<std::unique_ptr<GroupRun>>* keep = nullptr) {
  if (ngroups <= 0 || !groups) return fail(KH_BAD_ARGUMENT, "no groups");
  int32_t cpb = (opt && opt->circuits_per_wave > 0) ? opt->circuits_per_wave : 1;
  if (cpb > kBlock) return fail(KH_BAD_ARGUMENT, "circuits_per_wave above 64");
  std::unique_lock<std::mutex> lk(g_run_mu, std::defer_lock);
  if (!keep) lk.lock();
  std::vector<std::unique_ptr<GroupRun>> runs;
  Runner R;
  R.cpb = cpb;
  R.lds_staging = opt && opt->staging != 0;
  if (R.lds_staging) {
    int dev = 0, optin = 0, percu = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > 0)
        R.lds_limit = std::max<int64_t>(R.lds_limit, optin);
      if (hipDeviceGetAttribute(&percu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess)
        R.lds_per_cu = percu;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) R.ncu = ncu;
    }
  }
  R.waves = (opt && (opt->waves_per_simd == 2 || opt->waves_per_simd == 4)) ? opt->waves_per_simd : 1;
  R.wave = opt && opt->circuit_wave != 0;
  if (R.wave)
    return fail(KH_UNSUPPORTED,
                "circuit_wave is switched off on the GPU: its first device run did not finish and the cause is not "
                "known; its device code is checked on the CPU (tests/wave_host.cpp), bit for bit against k_persist's");
  for (int32_t g = 0; g < ngroups; g++) {
    const kh_group& G = groups[g];
    if (!G.topology || G.B <= 0 || !G.d_params) return fail(KH_BAD_ARGUMENT, "bad group " + std::to_string(g));
    auto r = std::make_unique<GroupRun>();
    r->topo = G.topology;
    r->B = G.B;
    r->gid = g;
    r->sw_points = sweep_points(p);
    r->want_sw_i = sw && sw->d_i;
    {  // MOSFET levels and internal nodes: every circuit must have circuit 0's
       // (the compiled currents and the unknown layout follow it)
      const Topology& T = G.topology->t;
      std::vector<double> lv((size_t)G.B);
      for (size_t k = 0; k < T.level_rows.size(); k++) {
        KH_HIP(hipMemcpy(lv.data(), G.d_params + (size_t)T.level_rows[k] * G.B, (size_t)G.B * 8,
                         hipMemcpyDeviceToHost));
        for (double v : lv)
          if (v != T.level_vals[k])
            return fail(KH_BAD_ARGUMENT, "group " + std::to_string(g) + ": MOSFET levels differ between circuits");
      }
      for (size_t k = 0; k < T.layout_rows.size(); k++) {
        KH_HIP(hipMemcpy(lv.data(), G.d_params + (size_t)T.layout_rows[k] * G.B, (size_t)G.B * 8,
                         hipMemcpyDeviceToHost));
        for (double v : lv)
          if ((v > 0) != (T.layout_vals[k] != 0))
            return fail(KH_BAD_ARGUMENT, "group " + std::to_string(g) +
                                             ": series resistances that decide internal nodes differ between circuits"
                                             " (parameter row " + std::to_string(T.layout_rows[k]) + ")");
      }
    }
    r->dense = use_dense(opt, G.topology->t.n);
    if (r->dense && G.topology->t.n > kMaxDense) return fail(KH_BAD_ARGUMENT, "dense backend above 4096 unknowns");
    kh_status s = r->setup(G.d_params, spec ? G.d_lane_spec : nullptr, G.d_nodeset);
    if (s != KH_OK) return s;
    R.groups.push_back(r.get());
    runs.push_back(std::move(r));
  }
  kh_status s = R.run(p);
  if (s != KH_OK) return s;
  for (int32_t g = 0; g < ngroups; g++) {
    const kh_group& G = groups[g];
    s = runs[g]->outputs(G.d_x, G.d_status, G.d_iters, G.d_accepted, G.d_rejected, G.d_ledger,
                         sw ? sw->d_x : nullptr, sw ? sw->d_iters : nullptr, sw ? sw->d_points : nullptr,
                         sw ? sw->d_i : nullptr);
    if (s != KH_OK) return s;
    s = runs[g]->report(p, G.d_report, G.d_report_ledger);
    if (s != KH_OK) return s;
  }
  KH_HIP(hipDeviceSynchronize());
  R.total();
  if (stats) *stats = R.stats;
  if (keep) *keep = std::move(runs);
  return KH_OK;
}

}  // namespace

extern "C" {

kh_status kh_batch_run(int32_t ngroups, const kh_group* groups, const kh_options* opt, const kh_tran_spec* spec,
                       kh_stats* stats) {
  CtlParams p = resolve(opt, spec ? AN_TRAN : ((opt && opt->polish) ? AN_OP_POLISH : AN_OP));
  if (spec) {
    kh_status s = tran_params(spec, p);
    if (s != KH_OK) return s;
  }
  return batch_run(ngroups, groups, opt, p, spec, nullptr, stats);
}

int32_t kh_candidate_size(const kh_candidate* c) {
  if (!c || !c->elems || c->n_elems <= 0 || !c->params || c->num_nodes < 1) return -KH_BAD_ARGUMENT;
  Topology T;
  if (!T.build(c->num_nodes, c->elems, c->n_elems, c->nparams, c->params)) return -KH_UNSUPPORTED;
  return T.n;
}

kh_status kh_generation_run(int32_t n, const kh_candidate* cands, const kh_options* opt, const kh_tran_spec* spec,
                            double* const* x_out, int32_t* status, int32_t* iters, int32_t* accepted,
                            int32_t* rejected, kh_stats* stats) {
  if (n <= 0 || !cands) return fail(KH_BAD_ARGUMENT, "no candidates");
  // Group by topology: node count, parameter count, element table, whether a
  // time scale is given, then the MOSFET levels (structural: they pick the
  // compiled currents) and the series resistances that create internal nodes
  // (structural: they decide the unknown layout), which the first candidate's
  // built topology locates.
  std::unordered_map<std::string, std::vector<int32_t>> by_key;
  std::vector<std::string> order;  // groups in order of first appearance
  for (int32_t i = 0; i < n; i++) {
    const kh_candidate& c = cands[i];
    if (!c.elems || c.n_elems <= 0 || !c.params || c.num_nodes < 1 || c.nparams < 0)
      return fail(KH_BAD_ARGUMENT, "candidate " + std::to_string(i) + ": bad arrays");
    std::string key(reinterpret_cast<const char*>(c.elems), (size_t)c.n_elems * sizeof(kh_element));
    const int32_t head[4] = {c.num_nodes, c.n_elems, c.nparams, (spec && c.lane_spec) ? 1 : 0};
    key.append(reinterpret_cast<const char*>(head), sizeof(head));
    auto f = by_key.find(key);
    if (f == by_key.end()) {
      order.push_back(key);
      by_key[key].push_back(i);
    } else {
      f->second.push_back(i);
    }
  }
  std::vector<std::vector<int32_t>> members;
  for (const auto& key : order) {
    const std::vector<int32_t>& idx = by_key[key];
    const kh_candidate& c0 = cands[idx[0]];
    Topology T;
    if (!T.build(c0.num_nodes, c0.elems, c0.n_elems, c0.nparams, c0.params))
      return fail(KH_UNSUPPORTED, "candidate " + std::to_string(idx[0]) + ": " + T.error);
    std::map<std::vector<double>, size_t> by_level;
    std::vector<std::vector<int32_t>> split;
    for (int32_t i : idx) {
      std::vector<double> lv;
      for (int32_t r : T.level_rows) lv.push_back(cands[i].params[r]);
      for (int32_t r : T.layout_rows) lv.push_back(cands[i].params[r] > 0 ? 1.0 : 0.0);
      auto f = by_level.find(lv);
      if (f == by_level.end()) {
        by_level[lv] = split.size();
        split.push_back({i});
      } else {
        split[f->second].push_back(i);
      }
    }
    for (auto& m : split) members.push_back(std::move(m));
  }
  // One topology and one set of device arrays per group.
  struct G {
    kh_topology t = nullptr;
    int32_t np = 0, nx = 0;
    DBuf par, lane, x, st, it, acc, rej;
    ~G() { if (t) kh_topology_destroy(t); }
  };
  std::vector<std::unique_ptr<G>> gs;
  std::vector<kh_group> groups;
  for (const auto& m : members) {
    const kh_candidate& c0 = cands[m[0]];
    auto g = std::make_unique<G>();
    kh_status s = kh_topology_create(c0.num_nodes, c0.elems, c0.n_elems, c0.nparams, c0.params, &g->t);
    if (s != KH_OK) return s;
    const int64_t B = (int64_t)m.size(), P = c0.nparams;
    g->np = (int32_t)P;
    g->nx = g->t->t.n;
    std::vector<double> hp((size_t)std::max<int64_t>(P, 1) * B);
    for (int64_t j = 0; j < B; j++)
      for (int64_t k = 0; k < P; k++) hp[(size_t)(k * B + j)] = cands[m[j]].params[k];
    KH_HIP(g->par.alloc(hp.size() * 8));
    KH_HIP(hipMemcpy(g->par.p, hp.data(), hp.size() * 8, hipMemcpyHostToDevice));
    const bool has_lane = spec && c0.lane_spec;
    if (has_lane) {
      std::vector<double> hl((size_t)4 * B);
      for (int64_t j = 0; j < B; j++)
        for (int k = 0; k < 4; k++) hl[(size_t)(k * B + j)] = cands[m[j]].lane_spec[k];
      KH_HIP(g->lane.alloc(hl.size() * 8));
      KH_HIP(hipMemcpy(g->lane.p, hl.data(), hl.size() * 8, hipMemcpyHostToDevice));
    }
    KH_HIP(g->x.alloc((size_t)g->nx * B * 8));
    KH_HIP(g->st.alloc((size_t)B * 4));
    KH_HIP(g->it.alloc((size_t)B * 4));
    KH_HIP(g->acc.alloc((size_t)B * 4));
    KH_HIP(g->rej.alloc((size_t)B * 4));
    kh_group k = {};
    k.topology = g->t;
    k.B = (int32_t)B;
    k.d_params = g->par.as<double>();
    k.d_lane_spec = has_lane ? g->lane.as<double>() : nullptr;
    k.d_x = g->x.as<double>();
    k.d_status = g->st.as<int32_t>();
    k.d_iters = g->it.as<int32_t>();
    k.d_accepted = g->acc.as<int32_t>();
    k.d_rejected = g->rej.as<int32_t>();
    groups.push_back(k);
    gs.push_back(std::move(g));
  }
  kh_status s = kh_batch_run((int32_t)groups.size(), groups.data(), opt, spec, stats);
  if (s != KH_OK) return s;
  // Results back to each candidate's own index.
  for (size_t gi = 0; gi < gs.size(); gi++) {
    const G& g = *gs[gi];
    const std::vector<int32_t>& m = members[gi];
    const int64_t B = (int64_t)m.size();
    std::vector<int32_t> hs(B), hi(B), ha(B), hr(B);
    std::vector<double> hx((size_t)g.nx * B);
    KH_HIP(hipMemcpy(hs.data(), g.st.p, (size_t)B * 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hi.data(), g.it.p, (size_t)B * 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(ha.data(), g.acc.p, (size_t)B * 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hr.data(), g.rej.p, (size_t)B * 4, hipMemcpyDeviceToHost));
    if (x_out) KH_HIP(hipMemcpy(hx.data(), g.x.p, hx.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < B; j++) {
      const int32_t i = m[j];
      if (status) status[i] = hs[j];
      if (iters) iters[i] = hi[j];
      if (accepted) accepted[i] = ha[j];
      if (rejected) rejected[i] = hr[j];
      if (x_out && x_out[i])
        for (int32_t k = 0; k < g.nx; k++) x_out[i][k] = hx[(size_t)k * B + j];
    }
  }
  return KH_OK;
}

kh_status kh_batch_dc(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                      const kh_dc_spec* spec, const double* d_nodeset, double* d_sweep_x, int32_t* d_status,
                      int32_t* d_points, int32_t* d_iters, double* d_ledger, kh_stats* stats) {
  kh_dc_outputs out = {};
  out.d_sweep_x = d_sweep_x;
  out.d_status = d_status;
  out.d_points = d_points;
  out.d_iters = d_iters;
  out.d_ledger = d_ledger;
  return kh_batch_dc_ex(t, B, d_params, opt, spec, d_nodeset, &out, stats);
}

kh_status kh_batch_dc_ex(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                         const kh_dc_spec* spec, const double* d_nodeset, const kh_dc_outputs* out,
                         kh_stats* stats) {
  if (!t || !spec || !out) return fail(KH_BAD_ARGUMENT, "no topology, sweep or outputs");
  if (spec->nsweeps < 1 || spec->nsweeps > 2) return fail(KH_BAD_ARGUMENT, "need 1 or 2 sweeps");
  CtlParams p = resolve(opt, AN_DC);
  p.swN = spec->nsweeps;
  int64_t total = 1;
  for (int s = 0; s < spec->nsweeps; s++) {
    // DCSweep.values, analysis/dc.go:31-46
    const double a = spec->start[s], b = spec->stop[s], h = spec->step[s];
    if (h == 0 || std::isnan(a + b + h) || std::isinf(a + b + h)) return fail(KH_BAD_ARGUMENT, "invalid DC sweep");
    const double cnt = (b - a) / h;
    if (cnt < 0 || cnt > 1000000) return fail(KH_BAD_ARGUMENT, "invalid DC sweep");
    p.swCount[s] = (int32_t)std::floor(cnt + 1e-9) + 1;
    p.swStart[s] = a;
    p.swStep[s] = h;
    const int32_t e = spec->element[s];
    if (e < 0 || e >= t->t.nel) return fail(KH_BAD_ARGUMENT, "swept element out of range");
    const int32_t kind = t->t.et[(size_t)e * ET_W + ET_KIND];
    if (kind != K_V && kind != K_I) return fail(KH_BAD_ARGUMENT, "swept element is not a V or I source");
    p.swElem[s] = e;
    total *= p.swCount[s];
  }
  if (total > 1000000) return fail(KH_BAD_ARGUMENT, "invalid DC sweep: too many points");
  kh_group g = {};
  g.topology = t; g.B = B; g.d_params = d_params; g.d_nodeset = d_nodeset; g.d_status = out->d_status;
  g.d_ledger = out->d_ledger;
  SweepOut sw;
  sw.d_x = out->d_sweep_x;
  sw.d_i = out->d_sweep_i;
  sw.d_iters = out->d_iters;
  sw.d_points = out->d_points;
  return batch_run(1, &g, opt, p, nullptr, &sw, stats);
}

kh_status kh_batch_ac(kh_topology t, int32_t B, const double* d_params, const kh_options* opt, const double* freqs,
                      int32_t nf, const double* d_nodeset, double* d_xac, int32_t* d_status, int32_t* d_points,
                      double* d_ledger, kh_stats* stats) {
  kh_ac_outputs out = {};
  out.d_xac = d_xac;
  out.d_status = d_status;
  out.d_points = d_points;
  out.d_ledger = d_ledger;
  return kh_batch_ac_ex(t, B, d_params, opt, freqs, nf, d_nodeset, &out, stats);
}

kh_status kh_batch_ac_ex(kh_topology t, int32_t B, const double* d_params, const kh_options* opt, const double* freqs,
                         int32_t nf, const double* d_nodeset, const kh_ac_outputs* out, kh_stats* stats) {
  if (!out) return fail(KH_BAD_ARGUMENT, "bad AC arguments");
  double* d_xac = out->d_xac;
  int32_t* d_status = out->d_status;
  int32_t* d_points = out->d_points;
  double* d_ledger = out->d_ledger;
  if (!t || B <= 0 || !d_params || !freqs || nf <= 0 || !d_xac || !d_status)
    return fail(KH_BAD_ARGUMENT, "bad AC arguments");
  for (int32_t k = 0; k < nf; k++)
    if (!(freqs[k] >= 0) || std::isinf(freqs[k])) return fail(KH_BAD_ARGUMENT, "bad AC frequency");
  std::lock_guard<std::mutex> lk(g_run_mu);
  // analysis.AC: the operating point, then polish (analysis/ac.go:85-90).
  CtlParams p = resolve(opt, AN_OP_POLISH);
  kh_group g = {};
  g.topology = t; g.B = B; g.d_params = d_params; g.d_nodeset = d_nodeset; g.d_status = d_status;
  std::vector<std::unique_ptr<GroupRun>> keep;
  kh_status s = batch_run(1, &g, opt, p, nullptr, nullptr, stats, &keep);
  if (s != KH_OK) return s;
  GroupRun& G = *keep[0];
  const Topology& T = t->t;
  const int n = T.n;
  // Frequencies in chunks of at most kAcLanes systems, fewer when the
  // factors of that many systems would not fit in device memory.
  constexpr int64_t kAcLanes = 1 << 18;
  DBuf dfreq, val, rhs, x, present, lst, lrr, pid, done, points, maxrr;
  KH_HIP(dfreq.alloc((size_t)nf * 8));
  KH_HIP(hipMemcpy(dfreq.p, freqs, (size_t)nf * 8, hipMemcpyHostToDevice));
  KH_HIP(val.alloc((size_t)T.nslots * B * sizeof(kcplx)));
  KH_HIP(rhs.alloc((size_t)n * B * sizeof(kcplx)));
  KH_HIP(present.alloc((size_t)B * T.nslots));
  KH_HIP(lst.alloc((size_t)B * 4));
  KH_HIP(done.alloc((size_t)B * 4));
  KH_HIP(points.alloc((size_t)B * 4));
  KH_HIP(maxrr.alloc((size_t)B * 8));
  DBuf mixl;
  DevPatterns pats;
  LuWorkT<kcplx> work;
  std::vector<int32_t> cpid(B, -1);
  AcArgs a = {};
  a.ckt = G.ctl_args(p).ckt;
  a.par = G.par.as<double>();
  a.st = G.st.as<double>();
  a.ci = G.ci.as<int32_t>();
  a.npar = T.nparams;
  a.nstate = std::max(T.nstate, 1);
  a.B = B;
  a.freqs = dfreq.as<double>();
  a.val = val.as<kcplx>();
  a.rhs = rhs.as<kcplx>();
  a.lstatus = lst.as<int32_t>();
  a.nf = nf;
  a.iout = out->d_iac;
  a.lout = out->d_ledger_f;
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto& e : ev) KH_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() { for (int k = 0; k < 2; k++) (void)hipEventDestroy(e[k]); }
  } guard{ev};
  auto t0 = std::chrono::steady_clock::now();
  // The AC pattern of each circuit (LoadAC skips zero admittances, so it
  // depends on the operating point and on whether omega is 0): the first
  // frequency is stamped once for every circuit to learn it, one ordering per
  // distinct pattern. The reference's CSparse keeps every entry it has seen
  // and re-orders when one appears (solver/sparse.go:114-143), so a sweep
  // that starts at 0 Hz has the 0 Hz pattern there and, from the first
  // nonzero frequency on, the union with that frequency's stamps (a jwC term
  // is zero only at omega = 0, so every nonzero frequency stamps the same
  // entries): that frequency is stamped too, into the same marks.
  int32_t fsplit = 0;  // frequencies below fsplit use cpid0, the rest cpid1
  if (nf > 1 && freqs[0] == 0) {
    fsplit = 1;
    while (fsplit < nf && freqs[fsplit] == 0) fsplit++;
    if (fsplit == nf) fsplit = 0;  // every frequency is 0 Hz: one pattern
  }
  std::vector<int32_t> cpid1(B, -1);
  a.BL = B;
  a.present = present.as<uint8_t>();
  std::vector<int32_t> wst(B);  // each circuit's operating-point status
  {
    std::vector<int32_t> hci((size_t)NCI * B);
    KH_HIP(hipMemcpy(hci.data(), G.ci.p, hci.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < B; c++) wst[c] = hci[(size_t)c * NCI + CI_STATUS];
    bool any_ok = false;
    for (int64_t c = 0; c < B; c++) any_ok = any_ok || wst[c] == ST_OK;
    if (!any_ok) nf = 0;  // no operating point succeeded: statuses are the OP's
    std::vector<uint8_t> pr((size_t)B * T.nslots);
    for (int probe = 0; probe < (fsplit > 0 ? 2 : 1); probe++) {
      a.f0 = probe == 0 ? 0 : fsplit;
      a.nfc = 1;
      hipLaunchKernelGGL(k_ac_stamp, dim3((unsigned)((B + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0, a);
      KH_HIP(hipGetLastError());
      KH_HIP(hipMemcpy(pr.data(), present.p, pr.size(), hipMemcpyDeviceToHost));
      std::vector<int32_t>& ids = probe == 0 ? cpid : cpid1;
      for (int64_t c = 0; c < B && any_ok && !G.dense; c++) {  // the dense backend has no pattern
        if (wst[c] != ST_OK) continue;
        std::vector<int32_t> slots;
        for (int s2 = 0; s2 < T.nslots; s2++)
          if (pr[(size_t)c * T.nslots + s2]) slots.push_back(s2);
        std::string key = slots_key(slots);
        auto f = pats.host.by_key.find(key);
        if (f != pats.host.by_key.end()) {
          ids[c] = f->second;
        } else {
          ids[c] = pats.host.add(analyze_pattern(n, T.slot_i, T.slot_j, slots), key);
        }
      }
    }
    if (fsplit == 0) cpid1 = cpid;
    KH_HIP(pats.sync());
  }
  const int64_t cap = G.dense ? (int64_t)n * n : lu_capacity(pats.host.max_fill, n);
  // Compiled schedules: FP64 sparse sweeps unless the caller opts out. Their
  // lanes hold values only (the structure is the schedule's), L and U sized
  // per round by the schedules in use (the chunk is sized for the structural
  // fill plus a margin); the lanes they leave get the per-lane LU's full
  // workspace, packed.
  const bool use_sched = !G.dense && p.precision == 0 && !(opt && opt->no_schedule);
  // L and U each hold about half the structural fill (|L| + |U| + n) under
  // diagonal pivoting; a margin of n each for pivots off the diagonal.
  const int64_t capS = pats.host.max_fill / 2 + (int64_t)n + 64;
  // Device bytes per lane: values, right-hand side, solution, status words and
  // the LU workspace (LuWorkT::ensure / ensure_dense, or the schedules' values).
  const int64_t full_lu_bytes = 3 * (int64_t)n * (int64_t)sizeof(kcplx) + 2 * cap * ((int64_t)sizeof(kcplx) + 4) +
                                (2 * cap + n) * 4 + 8 * ((int64_t)n + 1) + 24 * (int64_t)n + 8;
  const int64_t lane_bytes =
      ((int64_t)T.nslots + 2 * (int64_t)n) * (int64_t)sizeof(kcplx) + 16 +
      (G.dense ? cap * (int64_t)sizeof(kcplx) + 2 * (int64_t)n * (int64_t)sizeof(kcplx) + 4 * (int64_t)n + 4
       : use_sched ? 3 * (int64_t)n * (int64_t)sizeof(kcplx) + 2 * capS * (int64_t)sizeof(kcplx) + 16
                   : full_lu_bytes);
  size_t free_b = 0, total_b = 0;
  KH_HIP(hipMemGetInfo(&free_b, &total_b));
  const int64_t fit = std::max<int64_t>(1, (int64_t)(0.6 * (double)free_b) / (lane_bytes * B));
  int64_t max_lanes = kAcLanes;
  if (const char* e = std::getenv("KH_AC_CHUNK_LANES")) max_lanes = std::max<int64_t>(1, std::atoll(e));  // tests
  const int32_t fchunk =
      (int32_t)std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::max(nf, 1), max_lanes / B, fit}));
  const int64_t BLmax = (int64_t)fchunk * B;
  KH_HIP(val.alloc((size_t)T.nslots * BLmax * sizeof(kcplx)));
  KH_HIP(rhs.alloc((size_t)n * BLmax * sizeof(kcplx)));
  KH_HIP(x.alloc((size_t)n * BLmax * sizeof(kcplx)));
  KH_HIP(lst.alloc((size_t)BLmax * 4));
  KH_HIP(lrr.alloc((size_t)BLmax * 8));
  KH_HIP(pid.alloc((size_t)BLmax * 4));
  KH_HIP(mixl.alloc((size_t)BLmax * 4));
  std::vector<int32_t> hmix(p.precision ? (size_t)BLmax : 0);
  DevScheds scheds;
  DBuf lock, dep, dpiv, lustat, sx, sw, sud;
  if (use_sched && nf > 0) {
    KH_HIP(lock.alloc((size_t)BLmax * 4));
    KH_HIP(dep.alloc((size_t)BLmax * 4));
    KH_HIP(dpiv.alloc((size_t)BLmax * 4));
    KH_HIP(lustat.alloc((size_t)BLmax * 4));
    // blocked per workgroup (sched_lane_blocked): room for whole blocks of up to 64 lanes
    const size_t padded = (size_t)(BLmax + 63) / 64 * 64;
    KH_HIP(sx.alloc((size_t)n * padded * sizeof(kcplx)));
    KH_HIP(sw.alloc((size_t)n * padded * sizeof(kcplx)));
    KH_HIP(sud.alloc((size_t)n * padded * sizeof(kcplx)));
  }
  a.val = val.as<kcplx>();
  a.rhs = rhs.as<kcplx>();
  a.lstatus = lst.as<int32_t>();
  a.lrr = lrr.as<double>();
  a.present = nullptr;
  if (nf > 0) {
    if (G.dense) KH_HIP(work.ensure_dense(n, BLmax));
    else if (!use_sched) KH_HIP(work.ensure(n, BLmax, cap));
  }
  std::vector<int32_t> lp((size_t)BLmax), lp_dev;  // pattern id per lane (lp_dev: what the device holds)
  for (int32_t f0 = 0; f0 < nf; f0 += fchunk) {
    const int32_t nfc = std::min(fchunk, nf - f0);
    a.f0 = f0;
    a.nfc = nfc;
    a.BL = (int64_t)nfc * B;
    for (int64_t l = 0; l < a.BL; l++) lp[l] = std::max((f0 + l % nfc < fsplit ? cpid : cpid1)[l / nfc], 0);
    if (lp != lp_dev) {
      KH_HIP(hipMemcpy(pid.p, lp.data(), (size_t)a.BL * 4, hipMemcpyHostToDevice));
      lp_dev = lp;
    }
    // Per-lane kernels of the chunk in workgroups of kAcLaneBlock lanes: more
    // wavefronts in flight for their long per-lane loops (measured faster for
    // the per-lane LU as well: 13.8 s against 15.0 s on the 64×64 sweep).
    const int lane_block = kAcLaneBlock;
    const unsigned grid = (unsigned)((a.BL + lane_block - 1) / lane_block);
    hipLaunchKernelGGL(k_ac_stamp, dim3(grid), dim3(lane_block), 0, 0, a);
    KH_HIP(hipGetLastError());
    LuArgsT<kcplx> lu = {};
    lu.n = n;
    lu.nslots = T.nslots;
    lu.B = a.BL;
    lu.pats = pats.desc.as<PatDesc>();
    lu.pool = pats.pool.as<int32_t>();
    lu.pid = pid.as<int32_t>();
    lu.val = a.val;
    lu.rhs = a.rhs;
    lu.x = x.as<kcplx>();
    SchedChunk sc;
    double fallback_ms = 0;
    if (use_sched) {
      // Compiled schedules (sched_chunk) on value-only lanes; the lanes they
      // leave run the per-lane LU packed (k_ac_lane_lu) in pieces that fit.
      lu.xw = sx.as<kcplx>();
      lu.w = sw.as<kcplx>();
      lu.udiag = sud.as<kcplx>();
      std::vector<int64_t> left;
      if (kh_status ss = sched_chunk(a, lu, scheds, pats, lp, wst, n, T.nslots, cap, lock.as<int32_t>(),
                                     dep.as<int32_t>(), dpiv.as<int32_t>(), ev, sc, left);
          ss != KH_OK)
        return ss;
      size_t fb_free = 0, fb_total = 0;
      KH_HIP(hipMemGetInfo(&fb_free, &fb_total));
      const int64_t per = (int64_t)full_lu_bytes + ((int64_t)T.nslots + 2 * (int64_t)n) * (int64_t)sizeof(kcplx) + 12;
      const int64_t piece = std::max<int64_t>(1, (int64_t)(0.5 * (double)fb_free) / per);
      for (size_t at = 0; at < left.size(); at += (size_t)piece) {
        const int64_t m = std::min<int64_t>(piece, (int64_t)(left.size() - at));
        LuWorkT<kcplx> fw;
        DBuf dl, cval, crhs, cx, cpidb;
        KH_HIP(fw.ensure(n, m, cap));
        KH_HIP(dl.alloc((size_t)m * 8));
        KH_HIP(cval.alloc((size_t)T.nslots * m * sizeof(kcplx)));
        KH_HIP(crhs.alloc((size_t)n * m * sizeof(kcplx)));
        KH_HIP(cx.alloc((size_t)n * m * sizeof(kcplx)));
        KH_HIP(cpidb.alloc((size_t)m * 4));
        KH_HIP(hipMemcpy(dl.p, left.data() + at, (size_t)m * 8, hipMemcpyHostToDevice));
        std::vector<int32_t> cp((size_t)m);
        for (int64_t i = 0; i < m; i++) cp[i] = lp[left[at + i]];
        KH_HIP(hipMemcpy(cpidb.p, cp.data(), (size_t)m * 4, hipMemcpyHostToDevice));
        const int64_t tv = (int64_t)T.nslots * m, tr = (int64_t)n * m;
        hipLaunchKernelGGL(k_gather_soa, dim3((unsigned)((tv + 255) / 256)), dim3(256), 0, 0, a.val, a.BL,
                           (int64_t)T.nslots, dl.as<int64_t>(), m, cval.as<kcplx>());
        KH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_gather_soa, dim3((unsigned)((tr + 255) / 256)), dim3(256), 0, 0, a.rhs, a.BL, (int64_t)n,
                           dl.as<int64_t>(), m, crhs.as<kcplx>());
        KH_HIP(hipGetLastError());
        LuArgsT<kcplx> lc = {};
        lc.n = n;
        lc.nslots = T.nslots;
        lc.B = m;
        lc.pats = pats.desc.as<PatDesc>();
        lc.pool = pats.pool.as<int32_t>();
        lc.pid = cpidb.as<int32_t>();
        lc.val = cval.as<kcplx>();
        lc.rhs = crhs.as<kcplx>();
        lc.x = cx.as<kcplx>();
        fw.fill(lc);
        KH_HIP(hipEventRecord(ev[0], 0));
        hipLaunchKernelGGL(k_ac_lane_lu, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0, lc,
                           dl.as<int64_t>(), m, x.as<kcplx>(), a.BL, lustat.as<int32_t>());
        KH_HIP(hipGetLastError());
        KH_HIP(hipEventRecord(ev[1], 0));
        KH_HIP(hipEventSynchronize(ev[1]));
        float fms = 0;
        KH_HIP(hipEventElapsedTime(&fms, ev[0], ev[1]));
        fallback_ms += fms;
      }
    } else {
      work.fill(lu);
      if (!G.dense) KH_HIP(work.invalidate());
    }
    KH_HIP(hipEventRecord(ev[0], 0));
    hipLaunchKernelGGL(k_ac_solve, dim3(grid), dim3(lane_block), 0, 0, a, lu, p.ledgerTol, p.skipEnergy, d_xac,
                       p.precision, mixl.as<int32_t>(), use_sched ? lock.as<int32_t>() : nullptr,
                       use_sched ? lustat.as<int32_t>() : nullptr);
    KH_HIP(hipGetLastError());
    KH_HIP(hipEventRecord(ev[1], 0));
    hipLaunchKernelGGL(k_ac_reduce, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, 0, a, done.as<int32_t>(),
                       d_status, points.as<int32_t>(), maxrr.as<double>());
    KH_HIP(hipGetLastError());
    KH_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    KH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (stats) {
      // k_ac_solve: complex values and right-hand side read once, solution
      // written once, 16*(nnzA + 2n) bytes per system.
      int64_t nnz = G.dense ? (int64_t)n * n : 0;
      for (const auto& d : pats.host.desc) nnz = std::max<int64_t>(nnz, d.nnz);
      stats->solves += a.BL;
      stats->lu_seconds += ms * 1e-3;
      stats->lu_launches++;
      stats->bytes_algorithmic += a.BL * 16 * (nnz + 2 * (int64_t)n);
      stats->sched_solves += sc.ok;
      stats->sched_redos += sc.left;
      stats->lu_seconds += (sc.kernel_ms + fallback_ms) * 1e-3;
      stats->sched_probes += sc.probes;
      stats->sched_rounds += sc.rounds;
      stats->sched_host_seconds += sc.host_s;
      stats->sched_seconds += sc.kernel_ms * 1e-3;
      if (p.precision) {
        KH_HIP(hipMemcpy(hmix.data(), mixl.p, (size_t)a.BL * 4, hipMemcpyDeviceToHost));
        for (int64_t l = 0; l < a.BL; l++) {
          if (hmix[l] > 0) stats->refine_passes += hmix[l];
          if (hmix[l] < 0) stats->fp64_redos++;
        }
      }
    }
  }
  KH_HIP(hipDeviceSynchronize());
  if (d_points) KH_HIP(hipMemcpy(d_points, points.p, (size_t)B * 4, hipMemcpyDeviceToDevice));
  if (d_ledger) KH_HIP(hipMemcpy(d_ledger, maxrr.p, (size_t)B * 8, hipMemcpyDeviceToDevice));
  if (stats) stats->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return KH_OK;
}

kh_status kh_batch_op(kh_topology t, int32_t B, const double* d_params, const kh_options* opt, double* d_x,
                      int32_t* d_status, int32_t* d_iters, kh_stats* stats) {
  kh_group g = {};
  g.topology = t; g.B = B; g.d_params = d_params; g.d_x = d_x; g.d_status = d_status; g.d_iters = d_iters;
  return kh_batch_run(1, &g, opt, nullptr, stats);
}

kh_status kh_batch_tran(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                        const kh_tran_spec* spec, const double* d_lane_spec, double* d_x, int32_t* d_status,
                        int32_t* d_accepted, int32_t* d_rejected, double* d_ledger, kh_stats* stats) {
  if (!spec) return fail(KH_BAD_ARGUMENT, "no transient spec");
  kh_group g = {};
  g.topology = t; g.B = B; g.d_params = d_params; g.d_lane_spec = d_lane_spec; g.d_x = d_x; g.d_status = d_status;
  g.d_accepted = d_accepted; g.d_rejected = d_rejected; g.d_ledger = d_ledger;
  return kh_batch_run(1, &g, opt, spec, stats);
}

}  // extern "C"

// ---- JacobianTearing (kron/jacobian.go:66-712) ----------------------------------------------
namespace {

template <typename T>
__global__ void __launch_bounds__(kBlock) k_tear(TearPlanView P, int64_t B, int32_t refused, double reuse_tol,
                                                 const T* val, const T* rhs, T* x, T* tv, double* dv, int32_t* iv,
                                                 int64_t* cnt, int32_t* status, int32_t* any_fallback) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= B) return;
  const TearLane<T> L{P, B, c, val, rhs, x, tv, dv, iv, cnt, reuse_tol};
  const int st = L.solve(refused != 0);
  status[c] = st;
  if (st == TEAR_FALLBACK) atomicOr(any_fallback, 1);
}

// untorn (kron/jacobian.go:418-432): the refused systems solved by
// solver.NewMatrix(n, solver.Auto) from their stamps in slot order.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_tear_untorn(LuArgsT<T> a, int32_t dense, const int32_t* slot_i,
                                                        const int32_t* slot_j, int32_t* status) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= a.B || status[c] != TEAR_FALLBACK) return;
  int st;
  if (dense) {
    st = lu_dense<T>(a, c, slot_i, slot_j);
  } else {
    st = lu_factor<T>(a, c);
    if (st == LU_OK) st = lu_solve<T>(a, c);
  }
  status[c] = st;
}

template <typename T>
struct TearBatchT {
  int n = 0;
  int64_t B = 0;
  TearHost host;
  TearPlanView dview = {};
  DBuf plan_i, plan_p, tv, dv, iv, cnt, flag, status;
  // the untorn LU
  bool fb_dense = false, fb_ready = false;
  DBuf si, sj, pid;
  DevPatterns pats;
  Pattern pat;
  LuWorkT<T> work;
};

}  // namespace

struct kh_tear_s {
  std::vector<std::pair<int32_t, int32_t>> cuts;
  double reuse_tol = 0;
  kh_tear_stats st = {};
};

struct kh_tear_batch_s {
  kh_tear_s* jt = nullptr;
  int cplx = 0;
  std::unique_ptr<TearBatchT<double>> r;
  std::unique_ptr<TearBatchT<kcplx>> z;
};

struct kh_tmat_s {
  kh_tear_s* jt = nullptr;
  int n = 0;
  int cplx = 0;
  std::unordered_map<int64_t, int32_t> slot;
  std::vector<int32_t> ri, ci;
  std::vector<double> val;  // (re, im) pairs when complex
  kh_tear_batch_s* batch = nullptr;  // the plan and device state of the current pattern (null: pattern grew)
  DBuf dval, db, dx, dst;
  ~kh_tmat_s();
};

namespace {

template <typename T>
kh_status tear_batch_init(TearBatchT<T>& b, kh_tear_s* jt, int n, const std::vector<int32_t>& ri,
                          const std::vector<int32_t>& ci, int64_t B) {
  b.n = n;
  b.B = B;
  b.host = tear_plan(n, ri, ci, jt->cuts);
  const TearHost& h = b.host;
  if (!h.refused) jt->st.max_pieces = std::max<int64_t>(jt->st.max_pieces, (int64_t)h.pcs.size());
  // one int32 pool of every index array, and the pieces
  std::vector<int32_t> pool;
  std::vector<size_t> at;
  for (const std::vector<int32_t>* a : {&h.ri, &h.ci, &h.pieceOf, &h.local, &h.rowIdx, &h.off, &h.rows, &h.idx,
                                        &h.pslot, &h.pli, &h.plj, &h.ports}) {
    at.push_back(pool.size());
    pool.insert(pool.end(), a->begin(), a->end());
  }
  KH_HIP(b.plan_i.alloc(std::max<size_t>(1, pool.size()) * 4));
  if (!pool.empty()) KH_HIP(hipMemcpy(b.plan_i.p, pool.data(), pool.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(b.plan_p.alloc(std::max<size_t>(1, h.pcs.size()) * sizeof(TearPiece)));
  if (!h.pcs.empty())
    KH_HIP(hipMemcpy(b.plan_p.p, h.pcs.data(), h.pcs.size() * sizeof(TearPiece), hipMemcpyHostToDevice));
  TearPlanView v = h.v;
  const int32_t* base = b.plan_i.template as<int32_t>();
  v.ri = base + at[0]; v.ci = base + at[1]; v.pieceOf = base + at[2]; v.local = base + at[3];
  v.rowIdx = base + at[4]; v.off = base + at[5]; v.rows = base + at[6]; v.idx = base + at[7];
  v.pslot = base + at[8]; v.pli = base + at[9]; v.plj = base + at[10]; v.ports = base + at[11];
  v.pc = b.plan_p.template as<TearPiece>();
  b.dview = v;
  const int64_t nT = h.refused ? 0 : v.nT, nD = h.refused ? 0 : v.nD, nI = h.refused ? 0 : v.nI;
  KH_HIP(b.tv.alloc((size_t)nT * B * sizeof(T)));
  KH_HIP(b.dv.alloc((size_t)nD * B * 8));
  KH_HIP(b.iv.alloc((size_t)nI * B * 4));  // zero: no piece has factors yet
  KH_HIP(b.cnt.alloc((size_t)kTearCounters * B * 8));
  KH_HIP(b.flag.alloc(4));
  KH_HIP(b.status.alloc((size_t)B * 4));
  return KH_OK;
}

// The untorn LU's pattern and workspace, made when a system first needs it.
template <typename T>
kh_status tear_fallback_ready(TearBatchT<T>& b) {
  if (b.fb_ready) return KH_OK;
  const TearHost& h = b.host;
  b.fb_dense = h.n < 30;  // Backend.Resolve(Auto), solver/matrix.go:77-85
  if (b.fb_dense) {
    KH_HIP(b.si.alloc((size_t)std::max<int64_t>(1, h.nslots) * 4));
    KH_HIP(b.sj.alloc((size_t)std::max<int64_t>(1, h.nslots) * 4));
    if (h.nslots) {
      KH_HIP(hipMemcpy(b.si.p, h.ri.data(), (size_t)h.nslots * 4, hipMemcpyHostToDevice));
      KH_HIP(hipMemcpy(b.sj.p, h.ci.data(), (size_t)h.nslots * 4, hipMemcpyHostToDevice));
    }
    KH_HIP(b.work.ensure_dense(h.n, b.B));
  } else {
    std::vector<int32_t> slots(h.nslots);
    for (int64_t k = 0; k < h.nslots; k++) slots[k] = (int32_t)k;
    b.pat = analyze_pattern(h.n, h.ri, h.ci, slots);
    b.pats.host.add(b.pat, slots_key(slots));
    KH_HIP(b.pats.sync());
    KH_HIP(b.work.ensure(h.n, b.B, lu_capacity(b.pat.fill_estimate, h.n)));
    KH_HIP(b.pid.alloc((size_t)b.B * 4));
  }
  b.fb_ready = true;
  return KH_OK;
}

template <typename T>
kh_status tear_batch_solve(TearBatchT<T>& b, kh_tear_s* jt, const T* d_vals, const T* d_b, T* d_x, int32_t* d_status,
                           hipStream_t st) {
  const int64_t B = b.B;
  int32_t* status = d_status ? d_status : b.status.template as<int32_t>();
  KH_HIP(hipMemsetAsync(b.cnt.p, 0, (size_t)kTearCounters * B * 8, st));
  KH_HIP(hipMemsetAsync(b.flag.p, 0, 4, st));
  hipLaunchKernelGGL(k_tear<T>, dim3(grid_for(B)), dim3(kBlock), 0, st, b.dview, B, b.host.refused ? 1 : 0,
                     jt->reuse_tol, d_vals, d_b, d_x, b.tv.template as<T>(), b.dv.template as<double>(), b.iv.template as<int32_t>(),
                     b.cnt.template as<int64_t>(), status, b.flag.template as<int32_t>());
  KH_HIP(hipGetLastError());
  int32_t any = 0;
  KH_HIP(hipMemcpyAsync(&any, b.flag.p, 4, hipMemcpyDeviceToHost, st));
  KH_HIP(hipStreamSynchronize(st));
  if (any) {
    kh_status e = tear_fallback_ready(b);
    if (e != KH_OK) return e;
    LuArgsT<T> a = {};
    a.n = b.n;
    a.nslots = (int32_t)b.host.nslots;
    a.B = B;
    a.val = d_vals;
    a.rhs = d_b;
    a.x = d_x;
    if (!b.fb_dense) {
      a.pats = b.pats.desc.template as<PatDesc>();
      a.pool = b.pats.pool.template as<int32_t>();
      a.pid = b.pid.template as<int32_t>();
      b.work.fill(a);
    } else {
      a.lx = b.work.lx.template as<T>();
      a.udiag = b.work.udiag.template as<T>();
      a.w = b.work.w.template as<T>();
      a.prow = b.work.prow.template as<int32_t>();
      a.err_step = b.work.err.template as<int32_t>();
    }
    hipLaunchKernelGGL(k_tear_untorn<T>, dim3(grid_for(B)), dim3(kBlock), 0, st, a, b.fb_dense ? 1 : 0,
                       b.si.template as<int32_t>(), b.sj.template as<int32_t>(), status);
    KH_HIP(hipGetLastError());
  }
  std::vector<int64_t> cnt((size_t)kTearCounters * B);
  KH_HIP(hipMemcpyAsync(cnt.data(), b.cnt.p, cnt.size() * 8, hipMemcpyDeviceToHost, st));
  KH_HIP(hipStreamSynchronize(st));
  int64_t sum[kTearCounters] = {};
  for (int k = 0; k < kTearCounters; k++)
    for (int64_t c = 0; c < B; c++) sum[k] += cnt[(size_t)k * B + c];
  jt->st.solves += sum[TC_SOLVES];
  jt->st.factors += sum[TC_FACTORS];
  jt->st.reuses += sum[TC_REUSES];
  jt->st.stale_reuses += sum[TC_STALE];
  jt->st.fallbacks += sum[TC_FALLBACKS];
  jt->st.refinements += sum[TC_REFINES];
  return KH_OK;
}

kh_status tear_batch_make(kh_tear_s* jt, int n, const std::vector<int32_t>& ri, const std::vector<int32_t>& ci,
                          int64_t B, int cplx, kh_tear_batch_s** out) {
  auto h = std::make_unique<kh_tear_batch_s>();
  h->jt = jt;
  h->cplx = cplx;
  kh_status e;
  if (cplx) {
    h->z = std::make_unique<TearBatchT<kcplx>>();
    e = tear_batch_init(*h->z, jt, n, ri, ci, B);
  } else {
    h->r = std::make_unique<TearBatchT<double>>();
    e = tear_batch_init(*h->r, jt, n, ri, ci, B);
  }
  if (e != KH_OK) return e;
  *out = h.release();
  return KH_OK;
}

kh_status tear_pieces(const TearHost& h, int32_t* npieces, int32_t* piece_of) {
  if (npieces) *npieces = h.refused ? -1 : (int32_t)h.pcs.size();
  if (piece_of) std::memcpy(piece_of, h.pieceOf.data(), (size_t)h.n * 4);
  return KH_OK;
}

}  // namespace

kh_tmat_s::~kh_tmat_s() { delete batch; }

extern "C" {

kh_status kh_tear_create(const int32_t* cuts, int32_t ncuts, double reuse_tol, kh_tear* out) {
  if (!out || ncuts < 0 || (ncuts > 0 && !cuts) || !(reuse_tol >= 0)) return fail(KH_BAD_ARGUMENT, "bad argument");
  auto t = std::make_unique<kh_tear_s>();
  for (int32_t k = 0; k < ncuts; k++) t->cuts.push_back({cuts[2 * k], cuts[2 * k + 1]});
  t->reuse_tol = reuse_tol;
  *out = t.release();
  return KH_OK;
}

kh_status kh_tear_get_stats(kh_tear t, kh_tear_stats* out) {
  if (!t || !out) return fail(KH_BAD_ARGUMENT, "null argument");
  *out = t->st;
  return KH_OK;
}

void kh_tear_destroy(kh_tear t) { delete t; }

kh_status kh_tmat_create(kh_tear t, int32_t n, int32_t complex_values, kh_tmat* out) {
  if (!t || !out) return fail(KH_BAD_ARGUMENT, "null argument");
  if (n <= 0) return fail(KH_BADSIZE, "size " + std::to_string(n));  // solver.ErrBadSize, kron/jacobian.go:224-226
  auto m = std::make_unique<kh_tmat_s>();
  m->jt = t;
  m->n = n;
  m->cplx = complex_values ? 1 : 0;
  *out = m.release();
  return KH_OK;
}

kh_status kh_tmat_clear(kh_tmat m) {
  if (!m) return fail(KH_BAD_ARGUMENT, "null handle");
  std::fill(m->val.begin(), m->val.end(), 0.0);
  return KH_OK;
}

kh_status kh_tmat_add_batch(kh_tmat m, const int32_t* i, const int32_t* j, const double* v, int64_t count) {
  if (!m || (count > 0 && (!i || !j || !v))) return fail(KH_BAD_ARGUMENT, "null argument");
  const int w = m->cplx ? 2 : 1;
  for (int64_t k = 0; k < count; k++) {  // tornMat.Add, kron/jacobian.go:238-253
    const int ii = i[k], jj = j[k];
    if (ii < 0 || jj < 0 || ii >= m->n || jj >= m->n) continue;
    const int64_t key = (int64_t)ii * m->n + jj;
    auto it = m->slot.find(key);
    int32_t s;
    if (it == m->slot.end()) {
      s = (int32_t)m->ri.size();
      m->slot[key] = s;
      m->ri.push_back(ii);
      m->ci.push_back(jj);
      for (int q = 0; q < w; q++) m->val.push_back(0.0);
      delete m->batch;  // a new entry: a new plan, pieces without factors
      m->batch = nullptr;
    } else {
      s = it->second;
    }
    for (int q = 0; q < w; q++) m->val[(size_t)s * w + q] = m->val[(size_t)s * w + q] + v[k * w + q];
  }
  return KH_OK;
}

kh_status kh_tmat_solve(kh_tmat m, const double* b, double* x) {
  if (!m || !b || !x) return fail(KH_BAD_ARGUMENT, "null argument");
  const int w = m->cplx ? 2 : 1;
  if (!m->batch) {
    kh_status e = tear_batch_make(m->jt, m->n, m->ri, m->ci, 1, m->cplx, &m->batch);
    if (e != KH_OK) return e;
    KH_HIP(m->dval.alloc(std::max<size_t>(1, m->val.size()) * 8));
    KH_HIP(m->db.alloc((size_t)m->n * w * 8));
    KH_HIP(m->dx.alloc((size_t)m->n * w * 8));
    KH_HIP(m->dst.alloc(4));
  }
  if (!m->val.empty()) KH_HIP(hipMemcpy(m->dval.p, m->val.data(), m->val.size() * 8, hipMemcpyHostToDevice));
  KH_HIP(hipMemcpy(m->db.p, b, (size_t)m->n * w * 8, hipMemcpyHostToDevice));
  kh_status e = kh_tear_batch_solve(m->batch, m->dval.as<double>(), m->db.as<double>(), m->dx.as<double>(),
                                    m->dst.as<int32_t>(), nullptr);
  if (e != KH_OK) return e;
  int32_t hs = 0;
  KH_HIP(hipMemcpy(&hs, m->dst.p, 4, hipMemcpyDeviceToHost));
  if (hs == LU_OK || hs == LU_NONFINITE) KH_HIP(hipMemcpy(x, m->dx.p, (size_t)m->n * w * 8, hipMemcpyDeviceToHost));
  if (hs == LU_OK) return KH_OK;
  if (hs == LU_SINGULAR) return fail(KH_SINGULAR, "solver: singular matrix");
  if (hs == LU_NONFINITE) return fail(KH_NONFINITE, "solver: non-finite value in system");
  return lu_status_error(hs);
}

kh_status kh_tmat_absproduct(kh_tmat m, const double* x, double* y) {
  if (!m || !x || !y) return fail(KH_BAD_ARGUMENT, "null argument");
  for (int i = 0; i < m->n; i++) y[i] = 0;  // tornMat.AbsProduct, kron/jacobian.go:255-262
  for (size_t s = 0; s < m->ri.size(); s++) {
    if (m->cplx) {
      const kcplx v = {m->val[2 * s], m->val[2 * s + 1]}, xv = {x[2 * m->ci[s]], x[2 * m->ci[s] + 1]};
      y[m->ri[s]] += lu_mag(v) * lu_mag(xv);
    } else {
      y[m->ri[s]] += std::fabs(m->val[s]) * std::fabs(x[m->ci[s]]);
    }
  }
  return KH_OK;
}

kh_status kh_tmat_pieces(kh_tmat m, int32_t* npieces, int32_t* piece_of) {
  if (!m) return fail(KH_BAD_ARGUMENT, "null handle");
  if (!m->batch) {
    if (npieces) *npieces = -1;
    return KH_OK;
  }
  return kh_tear_batch_pieces(m->batch, npieces, piece_of);
}

void kh_tmat_destroy(kh_tmat m) { delete m; }

kh_status kh_tear_batch_create(kh_tear t, int32_t n, int64_t nslots, const int32_t* ri, const int32_t* ci,
                               int32_t B, int32_t complex_values, kh_tear_batch* out) {
  if (!t || !out || B <= 0 || nslots < 0 || (nslots > 0 && (!ri || !ci))) return fail(KH_BAD_ARGUMENT, "bad argument");
  if (n <= 0) return fail(KH_BADSIZE, "size " + std::to_string(n));
  std::vector<int32_t> r(ri, ri + nslots), c(ci, ci + nslots);
  std::unordered_map<int64_t, int> seen;
  for (int64_t k = 0; k < nslots; k++) {
    if (r[k] < 0 || c[k] < 0 || r[k] >= n || c[k] >= n) return fail(KH_BAD_ARGUMENT, "entry out of range");
    if (!seen.emplace((int64_t)r[k] * n + c[k], 1).second) return fail(KH_BAD_ARGUMENT, "duplicate entry");
  }
  return tear_batch_make(t, n, r, c, B, complex_values ? 1 : 0, out);
}

kh_status kh_tear_batch_solve(kh_tear_batch h, const double* d_vals, const double* d_b, double* d_x,
                              int32_t* d_status, void* stream) {
  if (!h || !d_b || !d_x || (!d_vals && (h->cplx ? h->z->host.nslots : h->r->host.nslots) > 0))
    return fail(KH_BAD_ARGUMENT, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (h->cplx)
    return tear_batch_solve(*h->z, h->jt, reinterpret_cast<const kcplx*>(d_vals), reinterpret_cast<const kcplx*>(d_b),
                            reinterpret_cast<kcplx*>(d_x), d_status, st);
  return tear_batch_solve(*h->r, h->jt, d_vals, d_b, d_x, d_status, st);
}

kh_status kh_tear_batch_pieces(kh_tear_batch h, int32_t* npieces, int32_t* piece_of) {
  if (!h) return fail(KH_BAD_ARGUMENT, "null handle");
  return tear_pieces(h->cplx ? h->z->host : h->r->host, npieces, piece_of);
}

void kh_tear_batch_destroy(kh_tear_batch h) { delete h; }

}  // extern "C"
