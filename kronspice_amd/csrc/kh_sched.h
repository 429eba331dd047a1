// kh_sched.h — compiled Gilbert–Peierls schedules: one pattern's factorization
// structure, shared by every system of a batch whose pivots follow it, run in
// lockstep with one lane per system.
//
// For a fixed pattern and a fixed pivot row per column, everything
// solver/sparse.go:199-401 does apart from arithmetic is fixed: the reach of
// every column (its order and membership), which reach rows update which
// column, the L and U patterns, the solve's loops. A schedule records that
// structure once, from a probe system factored on the host with the per-lane
// code (kh_lu.h). sched_lu then runs factor and solve for one system along
// the schedule: the same values meet the same operations in the same order as
// lu_factor / lu_solve, so its results are bit-identical to theirs — as long
// as the system's own pivot rule picks the schedule's pivot row at every
// column. sched_lu evaluates that rule on the system's own values at every
// column (the reference's argmax with its tie rule, the column-relative
// singularity threshold and the KLU diagonal preference) and returns LU_UNSURE
// at the first column where it would pick another row or fail, and for the
// cases lu_factor reports (non-finite or all-zero values, a non-finite
// solution): the caller redoes that system with lu_factor / lu_solve.
//
// Why: the schedule's indices are the same for every lane of a wavefront, so
// they are scalar loads shared by the lanes, and with the batch stored
// structure-of-arrays (entry i of system c at [i * B + c]) every value access
// is one coalesced transaction. The per-lane code instead walks each lane's
// own reach lists and index arrays, a chain of dependent loads per update.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "kh_lu.h"

namespace kh {

// One U entry of column k in the reach order of the factorization: the
// pivotal row r (the row pivoted at column j), and L(:, j)'s entry range
// [l0, l1) — the update x -= L(:, j) * x[r] and the U value U(j, k) = x[r].
struct alignas(16) SchedU {
  int32_t r, l0, l1, j;
};

// Schedule arrays are read-only while a kernel runs; on the device they are
// reached through the constant address space, so the compiler loads them with
// scalar instructions (one load per wavefront, into scalar registers) whenever
// the index is uniform — and it always is.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using SchedPtr = const __attribute__((address_space(4))) T*;
#else
template <class T> using SchedPtr = const T*;
#endif

// Device view of one schedule (all pointers device memory).
struct SchedView {
  int32_t n = 0;
  int32_t pid = -1;                   // the pattern it belongs to
  SchedPtr<int32_t> prow = nullptr;   // [n] pivot row of each column
  SchedPtr<int32_t> lp = nullptr;     // [n+1] L column starts
  SchedPtr<int32_t> li = nullptr;     // [nnzL] L rows (permuted ids), per column in reach order
  SchedPtr<int32_t> up = nullptr;     // [n+1] U column starts
  SchedPtr<SchedU> ue = nullptr;      // [nnzU]
  SchedPtr<int32_t> dc = nullptr;     // [n] 1: row k is a pivot candidate of column k (diagonal preference)
  // The compact working column (sched_lu_lds): column k's reach rows get slots
  // 0..m_k-1 — its U rows in reach order, its pivot row, its L rows in order.
  int32_t mmax = 0;                   // largest m_k
  SchedPtr<int32_t> tbase = nullptr;  // [n+1] start of column k's update targets in tpos
  SchedPtr<int32_t> tpos = nullptr;   // [updates] slot of each update's row, in the order of the updates
  SchedPtr<int32_t> apos = nullptr;   // [nnzA] slot of each A entry's row in its column
  SchedPtr<int32_t> dslot = nullptr;  // [n] slot of row k in column k when it is a pivot candidate, else -1
  SchedPtr<int32_t> ap = nullptr;     // the pattern's permuted CSC (PatDesc) and permutation
  SchedPtr<int32_t> arow = nullptr;
  SchedPtr<int32_t> aslot = nullptr;
  SchedPtr<int32_t> q = nullptr;
  SchedPtr<int32_t> qinv = nullptr;
};

// Host form of a schedule: the arrays SchedView points at, built by
// build_schedule from a probe factorization.
struct SchedHost {
  int32_t n = 0, pid = -1, mmax = 0;
  std::vector<int32_t> prow, lp, li, up, dc, tbase, tpos, apos, dslot;
  std::vector<SchedU> ue;
};

// Factors the probe system `val` (nslots values of pattern pid of pool pp) on
// the host with lu_factor and records its structure. Returns false when the
// probe does not factor (the schedule would be of no use: its systems take
// the per-lane path).
template <typename T>
inline bool build_schedule(const std::vector<PatDesc>& desc, const std::vector<int32_t>& pool, int32_t pid, int n,
                           int32_t nslots, const T* val, int64_t cap, SchedHost& out) {
  std::vector<T> xw(n), w(n), udiag(n), lx(cap), ux(cap), rhs(n), x(n);
  std::vector<int32_t> li(cap), ui(cap), lp(n + 1), up(n + 1), pinv(n), prow(n), xi(n), pstack(n), mark(n),
      rch(2 * cap + n), rp(n + 1), dk(n);
  int32_t err = 0, cpid = -1, pidv = pid;
  LuArgsT<T> a = {};
  a.n = n;
  a.nslots = nslots;
  a.B = 1;
  a.capL = a.capU = cap;
  a.capR = 2 * cap + n;
  a.pats = desc.data();
  a.pool = pool.data();
  a.pid = &pidv;
  a.val = val;
  a.rhs = rhs.data();
  a.x = x.data();
  a.xw = xw.data(); a.w = w.data(); a.udiag = udiag.data();
  a.lx = lx.data(); a.ux = ux.data(); a.li = li.data(); a.ui = ui.data();
  a.lp = lp.data(); a.up = up.data(); a.pinv = pinv.data(); a.prow = prow.data();
  a.xi = xi.data(); a.pstack = pstack.data(); a.mark = mark.data(); a.err_step = &err;
  a.rch = rch.data(); a.rp = rp.data(); a.dk = dk.data(); a.cpid = &cpid;
  if (lu_factor<T, true>(a, 0) != LU_OK) return false;
  const int64_t nl = lp[n], nu = up[n];
  out.n = n;
  out.pid = pid;
  out.prow.assign(prow.begin(), prow.end());
  out.lp.assign(lp.begin(), lp.end());
  out.li.assign(li.begin(), li.begin() + nl);
  out.up.assign(up.begin(), up.end());
  out.ue.resize(nu);
  for (int64_t q = 0; q < nu; q++) {
    const int32_t j = ui[q];
    out.ue[q] = SchedU{prow[j], lp[j], lp[j + 1], j};
  }
  // Row k is a candidate of column k when it is in the column's reach and was
  // not pivoted before column k (pinv[k] >= k once the factorization is done).
  out.dc.assign(n, 0);
  for (int k = 0; k < n; k++) out.dc[k] = (dk[k] != 0 && pinv[k] >= k) ? 1 : 0;
  // Compact slots (sched_lu_lds).
  const PatDesc pd = desc[pid];
  const int32_t* ap = pool.data() + pd.ap;
  const int32_t* arow = pool.data() + pd.arow;
  std::vector<int32_t> slot(n, -1);
  out.tbase.assign(n + 1, 0);
  out.apos.assign(pd.nnz, -1);
  out.dslot.assign(n, -1);
  out.tpos.clear();
  out.mmax = 0;
  for (int k = 0; k < n; k++) {
    const int u0 = up[k], u1 = up[k + 1], l0 = lp[k], l1 = lp[k + 1], P = u1 - u0;
    for (int q = u0; q < u1; q++) slot[prow[ui[q]]] = q - u0;
    slot[prow[k]] = P;
    for (int t = l0; t < l1; t++) slot[li[t]] = P + 1 + (t - l0);
    out.mmax = std::max(out.mmax, P + 1 + (l1 - l0));
    if (out.dc[k]) out.dslot[k] = slot[k];
    bool ok = !out.dc[k] || slot[k] >= 0;
    for (int p = ap[k]; p < ap[k + 1]; p++) {
      out.apos[p] = slot[arow[p]];
      ok = ok && out.apos[p] >= 0;
    }
    out.tbase[k] = (int32_t)out.tpos.size();
    for (int q = u0; q < u1; q++) {
      const int j = ui[q];
      for (int t = lp[j]; t < lp[j + 1]; t++) {
        out.tpos.push_back(slot[li[t]]);
        ok = ok && slot[li[t]] >= 0;
      }
    }
    if (!ok) return false;  // not a reach set (cannot happen for a factorization lu_factor produced)
    for (int q = u0; q < u1; q++) slot[prow[ui[q]]] = -1;
    slot[prow[k]] = -1;
    for (int t = l0; t < l1; t++) slot[li[t]] = -1;
  }
  out.tbase[n] = (int32_t)out.tpos.size();
  return true;
}

// A system's private arrays in the lockstep path: the working column (and the
// solve's y), the permuted right-hand side w, the U diagonal, and the L and U
// values, each with element i at [i * st]. Structure-of-arrays over the batch
// (sched_lane_soa) or blocked per workgroup (sched_lane_blocked: a workgroup's
// W lanes interleaved, one block after another, so an L column's entries of
// one workgroup are consecutive in memory).
template <typename T>
struct SchedLane {
  T *xw, *w, *udiag, *lx, *ux;
  int64_t st;
};
template <typename T>
KH_HD inline SchedLane<T> sched_lane_soa(const LuArgsT<T>& a, int64_t c) {
  return SchedLane<T>{a.xw + c, a.w + c, a.udiag + c, a.lx + c, a.ux + c, a.B};
}
template <typename T>
KH_HD inline SchedLane<T> sched_lane_blocked(const LuArgsT<T>& a, int64_t c, int W) {
  const int64_t blk = c / W, wl = c % W;
  return SchedLane<T>{a.xw + blk * a.n * W + wl, a.w + blk * a.n * W + wl, a.udiag + blk * a.n * W + wl,
                      a.lx + blk * a.capL * W + wl, a.ux + blk * a.capU * W + wl, W};
}

// The reference's first checks (solver/sparse.go:204-215): every value
// finite, not all of them zero (the largest magnitude is zero exactly when
// every value is zero, so no magnitude is needed).
template <typename T>
KH_HD inline bool sched_values_ok(const LuArgsT<T>& a, int64_t c) {
  const int64_t B = a.B;
  bool any = false;
  int sl = 0;
  for (; sl + kLuChunk <= a.nslots; sl += kLuChunk) {
    T v[kLuChunk];
#pragma unroll
    for (int u = 0; u < kLuChunk; u++) v[u] = a.val[(int64_t)(sl + u) * B + c];
#pragma unroll
    for (int u = 0; u < kLuChunk; u++) {
      if (!lu_finite(v[u])) return false;
      any = any || !lu_zero(v[u]);
    }
  }
  for (; sl < a.nslots; sl++) {
    const T v = a.val[(int64_t)sl * B + c];
    if (!lu_finite(v)) return false;
    any = any || !lu_zero(v);
  }
  return any;
}

// Entries handled together by the update loops: their loads are issued
// before any of them is used (the rows of one L column are distinct).
constexpr int kSchedChunk = 4;
constexpr int kSchedBatch = 8;  // L-column updates per batch

// solve of system c with the factors sched_lu / sched_lu_lds left in its
// lx, ux and udiag (solver/sparse.go:355-401): w = b permuted, L y = P w,
// U z = y, x = Q z. A non-finite entry returns LU_UNSURE.
template <typename T>
KH_HD inline int sched_solve(const SchedView& s, const LuArgsT<T>& a, int64_t c, const SchedLane<T>& P) {
  const int64_t B = a.B, st = P.st;
  const int n = s.n;
  T* const lx = P.lx;
  T* const ux = P.ux;
  T* const w = P.w;
  T* const y = P.xw;
  {
    int i = 0;
    for (; i + kLuChunk <= n; i += kLuChunk) {
      T v[kLuChunk];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) v[u] = a.rhs[(int64_t)(i + u) * B + c];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) w[(int64_t)s.qinv[i + u] * st] = v[u];
    }
    for (; i < n; i++) w[(int64_t)s.qinv[i] * st] = a.rhs[(int64_t)i * B + c];
  }
  for (int k = 0; k < n; k++) {
    const T yk = w[(int64_t)s.prow[k] * st];
    y[(int64_t)k * st] = yk;
    if (lu_zero(yk)) continue;
    const int l1 = s.lp[k + 1];
    for (int t = s.lp[k]; t < l1; t += kSchedBatch) {
      const int cnt = l1 - t;
      int64_t ix[kSchedBatch];
      T lv[kSchedBatch], wv[kSchedBatch];
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++)
        if (u < cnt) {
          ix[u] = (int64_t)s.li[t + u] * st;
          lv[u] = lx[(int64_t)(t + u) * st];
          wv[u] = w[ix[u]];
        }
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++)
        if (u < cnt) {
          const T prod = lv[u] * yk;
          w[ix[u]] = wv[u] - prod;
        }
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    const T zk = lu_div(y[(int64_t)k * st], P.udiag[(int64_t)k * st]);
    y[(int64_t)k * st] = zk;
    if (lu_zero(zk)) continue;
    const int u1 = s.up[k + 1];
    for (int qq = s.up[k]; qq < u1; qq += kSchedBatch) {
      const int cnt = u1 - qq;
      int64_t ix[kSchedBatch];
      T uv[kSchedBatch], yv[kSchedBatch];
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++)
        if (u < cnt) {
          ix[u] = (int64_t)s.ue[qq + u].j * st;
          uv[u] = ux[(int64_t)(qq + u) * st];
          yv[u] = y[ix[u]];
        }
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++)
        if (u < cnt) {
          const T prod = uv[u] * zk;
          y[ix[u]] = yv[u] - prod;
        }
    }
  }
  // A non-finite solution entry: the per-lane path writes the entries before
  // it and reports it, as the reference does.
  for (int k = 0; k < n; k++)
    if (!lu_finite(y[(int64_t)k * st])) return LU_UNSURE;
  {
    int k = 0;
    for (; k + kLuChunk <= n; k += kLuChunk) {
      T v[kLuChunk];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) v[u] = y[(int64_t)(k + u) * st];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) a.x[(int64_t)s.q[k + u] * B + c] = v[u];
    }
    for (; k < n; k++) a.x[(int64_t)s.q[k] * B + c] = y[(int64_t)k * st];
  }
  return LU_OK;
}

// factor + solve of system c along schedule s (see the file comment). The
// schedule's pattern must be the system's. Writes x on LU_OK; LU_UNSURE
// leaves the system to lu_factor / lu_solve.
// dep (optional) receives the column at which the system's own pivot rule
// departed from the schedule, or -2 when another check refused it; dpiv the
// row that rule picked there.
template <typename T>
KH_HD inline int sched_lu(const SchedView& s, const LuArgsT<T>& a, int64_t c, const SchedLane<T>& P,
                          int* dep = nullptr, int* dpiv = nullptr) {
  if (dep) *dep = -2;
  const int64_t B = a.B;
  const int n = s.n;
  // The reference's first checks (solver/sparse.go:204-215): a non-finite
  // value or an all-zero matrix is reported by the per-lane path.
  if (!sched_values_ok(a, c)) return LU_UNSURE;
  const int64_t st = P.st;
  T* const x = P.xw;  // x[i * st]: the working column of this system
  T* const lx = P.lx;
  T* const ux = P.ux;
  const T zero = lu_zero_value<T>();
  for (int k = 0; k < n; k++) {
    const int u0 = s.up[k], u1 = s.up[k + 1], l0 = s.lp[k], l1 = s.lp[k + 1];
    const int pk = s.prow[k];
    // x = 0 on the column's reach: its U rows, its pivot row, its L rows.
    for (int qq = u0; qq < u1; qq++) x[(int64_t)s.ue[qq].r * st] = zero;
    x[(int64_t)pk * st] = zero;
    for (int t = l0; t < l1; t++) x[(int64_t)s.li[t] * st] = zero;
    // Scatter A(:, k); its largest magnitude sets the singularity threshold.
    double colMax = 0.0;
    {
      int p = s.ap[k];
      const int pe = s.ap[k + 1];
      for (; p + kSchedChunk <= pe; p += kSchedChunk) {
        T v[kSchedChunk];
#pragma unroll
        for (int u = 0; u < kSchedChunk; u++) v[u] = a.val[(int64_t)s.aslot[p + u] * B + c];
#pragma unroll
        for (int u = 0; u < kSchedChunk; u++) {
          x[(int64_t)s.arow[p + u] * st] = v[u];
          const double av = lu_mag(v[u]);
          if (av > colMax) colMax = av;
        }
      }
      for (; p < pe; p++) {
        const T v = a.val[(int64_t)s.aslot[p] * B + c];
        x[(int64_t)s.arow[p] * st] = v;
        const double av = lu_mag(v);
        if (av > colMax) colMax = av;
      }
    }
    const double tiny = colMax * (double)n * 1e-14;
    // x = L \ A(:, k): pivotal reach rows in reach order; each one's x is
    // final when it is reached, and it is also U(j, k).
    for (int qq = u0; qq < u1; qq++) {
      const SchedU e = s.ue[qq];
      const T xj = x[(int64_t)e.r * st];
      ux[(int64_t)qq * st] = xj;
      if (lu_zero(xj)) continue;
      // L(:, j) in batches: every load of a batch is issued before any
      // update (the batch's rows are distinct), the tail masked by a uniform
      // count instead of a one-entry loop.
      for (int t = e.l0; t < e.l1; t += kSchedBatch) {
        const int cnt = e.l1 - t;
        int64_t ix[kSchedBatch];
        T lv[kSchedBatch], xv[kSchedBatch];
#pragma unroll
        for (int u = 0; u < kSchedBatch; u++)
          if (u < cnt) {
            ix[u] = (int64_t)s.li[t + u] * st;
            lv[u] = lx[(int64_t)(t + u) * st];
            xv[u] = x[ix[u]];
          }
#pragma unroll
        for (int u = 0; u < kSchedBatch; u++)
          if (u < cnt) {
            const T prod = lv[u] * xj;
            x[ix[u]] = xv[u] - prod;
          }
      }
    }
    // The pivot rule on this system's values (solver/sparse.go:254-277): the
    // candidates are the rows not yet pivotal, the schedule's pivot and L rows.
    double best = lu_mag(x[(int64_t)pk * st]);
    int piv = (best >= 0) ? pk : -1;  // a NaN magnitude is never taken
    if (piv < 0) best = -1.0;
    for (int t = l0; t < l1; t += kSchedBatch) {
      const int cnt = l1 - t;
      T xv[kSchedBatch];
      int rr[kSchedBatch];
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++)
        if (u < cnt) {
          rr[u] = s.li[t + u];
          xv[u] = x[(int64_t)rr[u] * st];
        }
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++)
        if (u < cnt) {
          const double av = lu_mag(xv[u]);
          if (av > best || (av == best && rr[u] < piv)) {
            best = av;
            piv = rr[u];
          }
        }
    }
    if (piv < 0 || best <= tiny) return LU_UNSURE;  // singular here: the per-lane path reports it (dep -2)
    if (s.dc[k]) {
      const double ak = lu_mag(x[(int64_t)k * st]);
      if (ak >= 1e-3 * best && ak > tiny) piv = k;
    }
    if (piv != pk) {
      if (dep) *dep = k;
      if (dpiv) *dpiv = piv;
      return LU_UNSURE;
    }
    const T pv = x[(int64_t)pk * st];
    P.udiag[(int64_t)k * st] = pv;
    {
      int t = l0;
      for (; t + kSchedChunk <= l1; t += kSchedChunk) {
        T xv[kSchedChunk];
#pragma unroll
        for (int u = 0; u < kSchedChunk; u++) xv[u] = x[(int64_t)s.li[t + u] * st];
#pragma unroll
        for (int u = 0; u < kSchedChunk; u++) lx[(int64_t)(t + u) * st] = lu_div(xv[u], pv);
      }
      for (; t < l1; t++) lx[(int64_t)t * st] = lu_div(x[(int64_t)s.li[t] * st], pv);
    }
  }
  return sched_solve(s, a, c, P);
}


// Element i of a compact working-column block XP (a pointer to doubles: a
// complex value is two of them), read and written by value.
template <typename XP> KH_HD inline double xld(XP X, int i, const double*) { return X[i]; }
template <typename XP> KH_HD inline kcplx xld(XP X, int i, const kcplx*) { return kcplx{X[2 * i], X[2 * i + 1]}; }
template <typename XP> KH_HD inline void xst(XP X, int i, double v) { X[i] = v; }
template <typename XP> KH_HD inline void xst(XP X, int i, kcplx v) {
  X[2 * i] = v.re;
  X[2 * i + 1] = v.im;
}

// sched_lu with the working column held compact: column k's reach rows in
// slots 0..m_k-1 (its U rows in reach order, its pivot row, its L rows), at
// element slot * W + wl of X (xld / xst) — the workgroup's W lanes' working
// columns (in LDS on
// the device: every dependent access of the factorization is then an LDS
// access; the schedule's tpos / apos / dslot give each access its slot). Same
// operations in the same order as sched_lu, so the same results.
template <typename T, typename XP>
KH_HD inline int sched_lu_lds(const SchedView& s, const LuArgsT<T>& a, int64_t c, const SchedLane<T>& PL, XP X,
                              int W, int wl, int* dep = nullptr, int* dpiv = nullptr) {
  if (dep) *dep = -2;
  if (!sched_values_ok(a, c)) return LU_UNSURE;
  const int64_t B = a.B, st = PL.st;
  const int n = s.n;
  T* const lx = PL.lx;
  T* const ux = PL.ux;
  const T zero = lu_zero_value<T>();
  const T* const tg = nullptr;  // selects the xld overload
  for (int k = 0; k < n; k++) {
    const int u0 = s.up[k], u1 = s.up[k + 1], l0 = s.lp[k], l1 = s.lp[k + 1];
    const int P = u1 - u0, m = P + 1 + (l1 - l0);
    for (int sl = 0; sl < m; sl++) xst(X, sl * W + wl, zero);
    double colMax = 0.0;
    {
      const int pe = s.ap[k + 1];
      for (int p = s.ap[k]; p < pe; p += kSchedBatch) {
        const int cnt = pe - p;
        T v[kSchedBatch];
#pragma unroll
        for (int u = 0; u < kSchedBatch; u++)
          if (u < cnt) v[u] = a.val[(int64_t)s.aslot[p + u] * B + c];
#pragma unroll
        for (int u = 0; u < kSchedBatch; u++)
          if (u < cnt) {
            xst(X, s.apos[p + u] * W + wl, v[u]);
            const double av = lu_mag(v[u]);
            if (av > colMax) colMax = av;
          }
      }
    }
    const double tiny = colMax * (double)n * 1e-14;
    // The updates of U entry qq; its first batch of L values arrives in
    // `first`, loaded one entry ahead (the loads of a batch are unconditional,
    // clamped to the entry's range, so their number is fixed and the wait for
    // them leaves the next entry's loads in flight).
    int tb = s.tbase[k];
    auto fetch = [&](T (&dst)[kSchedBatch], const SchedU& e) {
      const int last = e.l1 > e.l0 ? e.l1 - 1 : (e.l0 > 0 ? e.l0 - 1 : 0);
#pragma unroll
      for (int u = 0; u < kSchedBatch; u++) dst[u] = lx[(int64_t)(e.l0 + u < last ? e.l0 + u : last) * st];
    };
    auto entry = [&](int qq, const SchedU& e, const T (&first)[kSchedBatch]) {
      const int len = e.l1 - e.l0;
      const T xj = xld(X, (qq - u0) * W + wl, tg);
      ux[(int64_t)qq * st] = xj;
      if (!lu_zero(xj)) {
        for (int t = 0; t < len; t += kSchedBatch) {
          const int cnt = len - t;
          int xs[kSchedBatch];
          T lv[kSchedBatch];
#pragma unroll
          for (int u = 0; u < kSchedBatch; u++)
            if (u < cnt) xs[u] = s.tpos[tb + t + u] * W + wl;
          if (t == 0) {
#pragma unroll
            for (int u = 0; u < kSchedBatch; u++) lv[u] = first[u];
          } else {
#pragma unroll
            for (int u = 0; u < kSchedBatch; u++)
              if (u < cnt) lv[u] = lx[(int64_t)(e.l0 + t + u) * st];
          }
#pragma unroll
          for (int u = 0; u < kSchedBatch; u++)
            if (u < cnt) {
              const T prod = lv[u] * xj;
              xst(X, xs[u], xld(X, xs[u], tg) - prod);
            }
        }
      }
      tb += len;
    };
    {
      T pa[kSchedBatch], pb[kSchedBatch];
      SchedU ea{}, eb{};
      if (u0 < u1) {
        ea = s.ue[u0];
        fetch(pa, ea);
      }
      for (int qq = u0; qq < u1; qq += 2) {
        if (qq + 1 < u1) {
          eb = s.ue[qq + 1];
          fetch(pb, eb);
        }
        entry(qq, ea, pa);
        if (qq + 1 >= u1) break;
        if (qq + 2 < u1) {
          ea = s.ue[qq + 2];
          fetch(pa, ea);
        }
        entry(qq + 1, eb, pb);
      }
    }
    // The pivot rule on this system's values (solver/sparse.go:254-277).
    const int pk = s.prow[k];
    double best = lu_mag(xld(X, P * W + wl, tg));
    int piv = (best >= 0) ? pk : -1;  // a NaN magnitude is never taken
    if (piv < 0) best = -1.0;
    for (int t = l0; t < l1; t++) {
      const int r = s.li[t];
      const double av = lu_mag(xld(X, (P + 1 + (t - l0)) * W + wl, tg));
      if (av > best || (av == best && r < piv)) {
        best = av;
        piv = r;
      }
    }
    if (piv < 0 || best <= tiny) return LU_UNSURE;  // singular here: the per-lane path reports it (dep -2)
    const int ds = s.dslot[k];
    if (ds >= 0) {
      const double ak = lu_mag(xld(X, ds * W + wl, tg));
      if (ak >= 1e-3 * best && ak > tiny) piv = k;
    }
    if (piv != pk) {
      if (dep) *dep = k;
      if (dpiv) *dpiv = piv;
      return LU_UNSURE;
    }
    const T pv = xld(X, P * W + wl, tg);
    PL.udiag[(int64_t)k * st] = pv;
    for (int t = l0; t < l1; t++) lx[(int64_t)t * st] = lu_div(xld(X, (P + 1 + (t - l0)) * W + wl, tg), pv);
  }
  return sched_solve(s, a, c, PL);
}

}  // namespace kh
