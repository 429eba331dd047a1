// kh_lu.h — per-lane Gilbert–Peierls LU for batches of MNA systems on gfx950.
//
// One lane (work-item) owns one circuit. All per-circuit arrays are stored
// structure-of-arrays with the circuit index fastest: entry i of circuit c is
// at base[i * B + c]. Lanes of a wavefront that share a sparsity pattern and
// a pivot sequence therefore touch the same i at the same time and every
// access is a coalesced 64-lane transaction; lanes whose pivots differ still
// compute correctly, only with scattered accesses.
//
// The numeric factorization follows kronspice's sparseLU.factor / reach /
// solve (solver/sparse.go:199-401) operation for operation — the same reach
// order, the same `x -= l*x_j` updates (no FMA contraction: the library is
// built with -ffp-contract=off), true divisions, the same pivot rule with
// column-relative singularity threshold and KLU diagonal preference — so
// pivot rows, L/U patterns and solutions are bit-identical to the reference
// for the same stamped values.
#pragma once
#include <stdint.h>
#include <math.h>

#ifndef KH_HD
#define KH_HD __host__ __device__
#endif

namespace kh {

// LU_UNSURE: the mixed-precision factorization or refinement could not
// certify its result (see lu_factor_mixed); the caller redoes the system in
// FP64. It never leaves the library.
enum LuStatus { LU_OK = 0, LU_SINGULAR = 1, LU_NONFINITE = 2, LU_CAPACITY = 14, LU_UNSURE = 16 };

// Complex values (CSparse, solver/sparse.go:495-531) with Go's complex128
// arithmetic: componentwise + and -, (ac - bd) + (ad + bc)i products, Smith's
// division (R. L. Smith, Algorithm 116, Commun. ACM 5(8):435, 1962) and
// magnitudes as math.Hypot computes them.
struct kcplx {
  double re, im;
};
KH_HD inline kcplx operator+(kcplx a, kcplx b) { return {a.re + b.re, a.im + b.im}; }
KH_HD inline kcplx operator-(kcplx a, kcplx b) { return {a.re - b.re, a.im - b.im}; }
KH_HD inline kcplx operator-(kcplx a) { return {-a.re, -a.im}; }
KH_HD inline kcplx operator*(kcplx a, kcplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
KH_HD inline kcplx conj(kcplx a) { return {a.re, -a.im}; }
KH_HD inline double go_hypot(double p, double q) {  // math.Hypot: p * sqrt(1 + (q/p)^2), p >= q
  p = fabs(p);
  q = fabs(q);
  if (isinf(p) || isinf(q)) return __builtin_huge_val();
  if (isnan(p) || isnan(q)) return __builtin_nan("");
  if (p < q) {
    double t = p;
    p = q;
    q = t;
  }
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}
KH_HD inline double lu_mag(double v) { return fabs(v); }
KH_HD inline double lu_mag(kcplx v) { return go_hypot(v.re, v.im); }
KH_HD inline bool lu_zero(double v) { return v == 0; }
KH_HD inline bool lu_zero(kcplx v) { return v.re == 0 && v.im == 0; }
KH_HD inline bool lu_finite(double v) { return !isnan(v) && !isinf(v); }
KH_HD inline bool lu_finite(kcplx v) { return lu_finite(v.re) && lu_finite(v.im); }
KH_HD inline double lu_div(double a, double b) { return a / b; }
KH_HD inline kcplx lu_div(kcplx n, kcplx m) {
  double e, f;
  if (fabs(m.re) >= fabs(m.im)) {
    const double ratio = m.im / m.re;
    const double denom = m.re + ratio * m.im;
    e = (n.re + n.im * ratio) / denom;
    f = (n.im - n.re * ratio) / denom;
  } else {
    const double ratio = m.re / m.im;
    const double denom = m.im + ratio * m.re;
    e = (n.re * ratio + n.im) / denom;
    f = (n.im * ratio - n.re) / denom;
  }
  return {e, f};
}
KH_HD inline double& lu_re(double& v) { return v; }
KH_HD inline double& lu_re(kcplx& v) { return v.re; }
template <typename T> KH_HD inline T lu_zero_value();
template <> KH_HD inline double lu_zero_value<double>() { return 0.0; }
template <> KH_HD inline kcplx lu_zero_value<kcplx>() { return {0.0, 0.0}; }

// Address-space policy of a lane's arrays: generic pointers, or (device
// compilation only) pointers into LDS, for a circuit whose working set is
// staged there, so its accesses compile to LDS instructions instead of flat
// ones (a flat access waits on the vector-memory counter, which also covers
// every store still in flight to HBM).
struct AsGen {
  template <class T> using P = T*;
};
struct AsLds {
#if defined(__HIP_DEVICE_COMPILE__)
  template <class T> using P = __attribute__((address_space(3))) T*;
#else
  template <class T> using P = T*;
#endif
};

// Permuted column structure of one pattern (solver/sparse.go:154-187):
// column k holds rows arow[ap[k]..ap[k+1]) (permuted ids, ascending) with
// values at slot aslot[...] of the lane's value array.
struct PatDesc {
  int32_t q;      // offset of q[n] in the pattern pool
  int32_t qinv;   // offset of qinv[n]
  int32_t ap;     // offset of ap[n+1]
  int32_t arow;   // offset of arow[nnz]
  int32_t aslot;  // offset of aslot[nnz]
  int32_t nnz;
  int32_t pad0, pad1;
};

// Everything a lane needs to factor and solve. Pointers are device memory.
// T = double (real systems) or kcplx (AC systems).
template <typename T, typename AS = AsGen>
struct LuArgsT {
  template <class U> using P = typename AS::template P<U>;
  int32_t n;          // unknowns
  int32_t nslots;     // value slots per lane
  int64_t B;          // lanes in the batch (SoA stride)
  int64_t capL, capU; // L and U entry capacity per lane
  const PatDesc* pats;
  const int32_t* pool;
  P<const int32_t> pid;       // [B] pattern id per lane
  P<const T> val;             // [nslots][B]
  P<const T> rhs;             // [n][B]
  P<T> x;                     // [n][B] solution (written on success)
  // workspace
  P<T> xw;                    // [n][B] dense working column
  P<T> w;                     // [n][B] permuted right-hand side
  P<T> udiag;                 // [n][B]
  P<T> lx;                    // [capL][B]
  P<T> ux;                    // [capU][B]
  P<int32_t> li;              // [capL][B]
  P<int32_t> ui;              // [capU][B]
  P<int32_t> lp;              // [n+1][B]
  P<int32_t> up;              // [n+1][B]
  P<int32_t> pinv;            // [n][B]
  P<int32_t> prow;            // [n][B]
  P<int32_t> xi;              // [n][B]
  P<int32_t> pstack;          // [n][B]
  P<int32_t> mark;            // [n][B]
  P<int32_t> err_step;        // [B] pivot step of a singular failure
  // reach cache of the previous factorization (see lu_factor)
  int64_t capR;             // entries of rch per lane (>= |L| + |U| + n)
  P<int32_t> rch;             // [capR][B] reach lists of all columns, concatenated
  P<int32_t> rp;              // [n+1][B] start of column k's reach in rch
  P<int32_t> dk;              // [n][B] 1 if row k is in the reach of column k
  P<int32_t> cpid;            // [B] pattern id the cache belongs to, -1: none
};
using LuArgs = LuArgsT<double>;

KH_HD inline bool kh_finite(double v) { return !isnan(v) && !isinf(v); }

constexpr int kLuChunk = 8;  // independent loads issued together (see lu_solve)

// reach, solver/sparse.go:301-352: non-recursive DFS from every row of
// A(:,k) through the graph of L; returns top, the reach is xi[top..n).
// mark[r] == k + 1 means "visited while processing column k" (the reference's
// generation counter, restarted per factorization).
template <typename T, bool Unit = false, typename AS = AsGen>
KH_HD inline int lu_reach(const LuArgsT<T, AS>& a, int64_t c, int k, const int32_t* ap, const int32_t* arow) {
  const int64_t B = Unit ? 1 : a.B;
  if (Unit) c = 0;
  const int n = a.n;
  const int32_t gen = k + 1;
  int top = n;
  for (int p = ap[k]; p < ap[k + 1]; p++) {
    int32_t r = arow[p];
    if (a.mark[(int64_t)r * B + c] == gen) continue;
    int head = 0;
    a.xi[c] = r;  // stack[0] = r (xi[0..) doubles as the DFS stack)
    while (head >= 0) {
      int32_t v = a.xi[(int64_t)head * B + c];
      int32_t j = a.pinv[(int64_t)v * B + c];
      if (a.mark[(int64_t)v * B + c] != gen) {
        a.mark[(int64_t)v * B + c] = gen;
        if (j >= 0) a.pstack[(int64_t)head * B + c] = a.lp[(int64_t)j * B + c];
      }
      bool done = true;
      if (j >= 0) {
        int32_t end = a.lp[(int64_t)(j + 1) * B + c];
        for (int32_t q = a.pstack[(int64_t)head * B + c]; q < end; q++) {
          int32_t wv = a.li[(int64_t)q * B + c];
          if (a.mark[(int64_t)wv * B + c] == gen) continue;
          a.pstack[(int64_t)head * B + c] = q + 1;
          head++;
          a.xi[(int64_t)head * B + c] = wv;
          done = false;
          break;
        }
      }
      if (done) {
        head--;
        top--;
        a.xi[(int64_t)top * B + c] = v;
      }
    }
  }
  return top;
}

// factor, solver/sparse.go:199-293. Unit = true: a single system stored with
// unit stride (one circuit's own arrays), B = 1 and c = 0 known at compile time.
//
// Reach reuse: the reach of column k (its order and membership) depends only
// on the pattern and on the pivot rows of columns < k. A circuit refactored
// with the same pattern and, so far, the same pivot rows as its previous
// factorization has the same reach lists, so they are replayed from the
// per-circuit cache rch/rp/dk instead of re-running the depth-first search.
// The pivot rule is still evaluated at every column exactly as the reference
// does; at the first column whose pivot differs, the search resumes. Results
// are identical to re-running the search every time.
template <typename T, bool Unit = false, typename AS = AsGen>
KH_HD inline int lu_factor(const LuArgsT<T, AS>& a, int64_t c) {
  const int64_t B = Unit ? 1 : a.B;
  if (Unit) c = 0;
  const int n = a.n;
  const int32_t pidc = a.pid[c];
  const PatDesc pd = a.pats[pidc];
  const int32_t* ap = a.pool + pd.ap;
  const int32_t* arow = a.pool + pd.arow;
  const int32_t* aslot = a.pool + pd.aslot;
  double scale = 0.0;
  {
    int s = 0;
    for (; s + kLuChunk <= a.nslots; s += kLuChunk) {
      T v[kLuChunk];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) v[u] = a.val[(int64_t)(s + u) * B + c];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) {
        if (!lu_finite(v[u])) return LU_NONFINITE;
        const double av = lu_mag(v[u]);
        if (av > scale) scale = av;
      }
    }
    for (; s < a.nslots; s++) {
      T v = a.val[(int64_t)s * B + c];
      if (!lu_finite(v)) return LU_NONFINITE;
      double av = lu_mag(v);
      if (av > scale) scale = av;
    }
  }
  if (scale == 0) {
    a.err_step[c] = -1;
    return LU_SINGULAR;
  }
  bool reuse = a.cpid[c] == pidc;
  a.cpid[c] = -1;  // the cache is valid again only once this factorization completes
  for (int i = 0; i < n; i++) a.pinv[(int64_t)i * B + c] = -1;
  if (!reuse)
    for (int i = 0; i < n; i++) a.mark[(int64_t)i * B + c] = 0;
  int64_t nl = 0, nu = 0, nr = 0;
  for (int k = 0; k < n; k++) {
    a.lp[(int64_t)k * B + c] = (int32_t)nl;
    a.up[(int64_t)k * B + c] = (int32_t)nu;
    int64_t r0 = nr, r1;
    bool diag_in_reach;
    if (reuse) {
      r1 = a.rp[(int64_t)(k + 1) * B + c];
      diag_in_reach = a.dk[(int64_t)k * B + c] != 0;
    } else {
      int top = lu_reach<T, Unit, AS>(a, c, k, ap, arow);
      r1 = r0 + (n - top);
      if (r1 > a.capR) return LU_CAPACITY;
      for (int p = top; p < n; p++) a.rch[(r0 + p - top) * B + c] = a.xi[(int64_t)p * B + c];
      diag_in_reach = a.mark[(int64_t)k * B + c] == k + 1;
      a.rp[(int64_t)(k + 1) * B + c] = (int32_t)r1;
      a.dk[(int64_t)k * B + c] = diag_in_reach ? 1 : 0;
    }
    nr = r1;
    {
      int64_t p = r0;
      for (; p + kLuChunk <= r1; p += kLuChunk) {
        int32_t rr[kLuChunk];
#pragma unroll
        for (int u = 0; u < kLuChunk; u++) rr[u] = a.rch[(p + u) * B + c];
#pragma unroll
        for (int u = 0; u < kLuChunk; u++) a.xw[(int64_t)rr[u] * B + c] = lu_zero_value<T>();
      }
      for (; p < r1; p++) a.xw[(int64_t)a.rch[p * B + c] * B + c] = lu_zero_value<T>();
    }
    double colMax = 0.0;
    {
      int p = ap[k];
      const int pe = ap[k + 1];
      for (; p + kLuChunk <= pe; p += kLuChunk) {
        T v[kLuChunk];
#pragma unroll
        for (int u = 0; u < kLuChunk; u++) v[u] = a.val[(int64_t)aslot[p + u] * B + c];
#pragma unroll
        for (int u = 0; u < kLuChunk; u++) {
          a.xw[(int64_t)arow[p + u] * B + c] = v[u];
          const double av = lu_mag(v[u]);
          if (av > colMax) colMax = av;
        }
      }
      for (; p < pe; p++) {
        T v = a.val[(int64_t)aslot[p] * B + c];
        a.xw[(int64_t)arow[p] * B + c] = v;
        double av = lu_mag(v);
        if (av > colMax) colMax = av;
      }
    }
    double tiny = colMax * (double)n * 1e-14;
    // x = L \ A(:,k) in reach order.
    for (int64_t p = r0; p < r1; p++) {
      int32_t r = a.rch[p * B + c];
      int32_t j = a.pinv[(int64_t)r * B + c];
      if (j < 0) continue;
      T xj = a.xw[(int64_t)r * B + c];
      if (lu_zero(xj)) continue;
      int32_t q1 = a.lp[(int64_t)(j + 1) * B + c];
      for (int32_t q = a.lp[(int64_t)j * B + c]; q < q1; q++) {
        int64_t ix = (int64_t)a.li[(int64_t)q * B + c] * B + c;
        T prod = a.lx[(int64_t)q * B + c] * xj;
        a.xw[ix] = a.xw[ix] - prod;
      }
    }
    // Pivot among the rows not yet pivotal; U entries in reach order.
    int32_t piv = -1;
    double best = -1.0;
    for (int64_t p = r0; p < r1; p++) {
      int32_t r = a.rch[p * B + c];
      int32_t pr = a.pinv[(int64_t)r * B + c];
      if (pr >= 0) {
        if (nu >= a.capU) return LU_CAPACITY;
        a.ui[nu * B + c] = pr;
        a.ux[nu * B + c] = a.xw[(int64_t)r * B + c];
        nu++;
        continue;
      }
      double av = lu_mag(a.xw[(int64_t)r * B + c]);
      if (av > best || (av == best && r < piv)) {
        best = av;
        piv = r;
      }
    }
    if (piv < 0 || best <= tiny) {
      a.err_step[c] = k;
      return LU_SINGULAR;
    }
    // KLU diagonal preference (solver/sparse.go:274-277).
    if (a.pinv[(int64_t)k * B + c] < 0 && diag_in_reach) {
      double ak = lu_mag(a.xw[(int64_t)k * B + c]);
      if (ak >= 1e-3 * best && ak > tiny) piv = k;
    }
    if (reuse && piv != a.prow[(int64_t)k * B + c]) {
      // Pivot row differs from the cached factorization: later reaches change.
      reuse = false;
      for (int i = 0; i < n; i++) a.mark[(int64_t)i * B + c] = 0;
    }
    T pv = a.xw[(int64_t)piv * B + c];
    a.udiag[(int64_t)k * B + c] = pv;
    a.pinv[(int64_t)piv * B + c] = k;
    a.prow[(int64_t)k * B + c] = piv;
    for (int64_t p = r0; p < r1; p++) {
      int32_t r = a.rch[p * B + c];
      if (a.pinv[(int64_t)r * B + c] < 0) {
        if (nl >= a.capL) return LU_CAPACITY;
        a.li[nl * B + c] = r;
        a.lx[nl * B + c] = lu_div(a.xw[(int64_t)r * B + c], pv);
        nl++;
      }
    }
  }
  a.lp[(int64_t)n * B + c] = (int32_t)nl;
  a.up[(int64_t)n * B + c] = (int32_t)nu;
  a.rp[c] = 0;
  a.cpid[c] = pidc;
  return LU_OK;
}

// solve, solver/sparse.go:355-401 (the factorization is held in the lane's
// workspace). x is written in original unknown order, stopping at the first
// non-finite entry as the reference does.
template <typename T, bool Unit = false, typename AS = AsGen>
KH_HD inline int lu_solve(const LuArgsT<T, AS>& a, int64_t c) {
  const int64_t B = Unit ? 1 : a.B;
  if (Unit) c = 0;
  const int n = a.n;
  const PatDesc pd = a.pats[a.pid[c]];
  const int32_t* q = a.pool + pd.q;
  const int32_t* qinv = a.pool + pd.qinv;
  {
    // The loads of kLuChunk entries are issued before any is used (one lane's
    // program is a chain of memory latencies); the same values move.
    int i = 0;
    for (; i + kLuChunk <= n; i += kLuChunk) {
      int32_t qi[kLuChunk];
      T v[kLuChunk];
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) {
        qi[u] = qinv[i + u];
        v[u] = a.rhs[(int64_t)(i + u) * B + c];
      }
#pragma unroll
      for (int u = 0; u < kLuChunk; u++) a.w[(int64_t)qi[u] * B + c] = v[u];
    }
    for (; i < n; i++) a.w[(int64_t)qinv[i] * B + c] = a.rhs[(int64_t)i * B + c];
  }
  const auto y = a.xw;
  for (int k = 0; k < n; k++) {
    T yk = a.w[(int64_t)a.prow[(int64_t)k * B + c] * B + c];
    y[(int64_t)k * B + c] = yk;
    if (lu_zero(yk)) continue;
    int32_t p1 = a.lp[(int64_t)(k + 1) * B + c];
    for (int32_t p = a.lp[(int64_t)k * B + c]; p < p1; p++) {
      int64_t ix = (int64_t)a.li[(int64_t)p * B + c] * B + c;
      T prod = a.lx[(int64_t)p * B + c] * yk;
      a.w[ix] = a.w[ix] - prod;
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    T zk = lu_div(y[(int64_t)k * B + c], a.udiag[(int64_t)k * B + c]);
    y[(int64_t)k * B + c] = zk;
    if (lu_zero(zk)) continue;
    int32_t p1 = a.up[(int64_t)(k + 1) * B + c];
    for (int32_t p = a.up[(int64_t)k * B + c]; p < p1; p++) {
      int64_t ix = (int64_t)a.ui[(int64_t)p * B + c] * B + c;
      T prod = a.ux[(int64_t)p * B + c] * zk;
      y[ix] = y[ix] - prod;
    }
  }
  int k = 0;
  for (; k + kLuChunk <= n; k += kLuChunk) {
    int32_t qk[kLuChunk];
    T v[kLuChunk];
#pragma unroll
    for (int u = 0; u < kLuChunk; u++) {
      qk[u] = q[k + u];
      v[u] = y[(int64_t)(k + u) * B + c];
    }
#pragma unroll
    for (int u = 0; u < kLuChunk; u++) {
      if (!lu_finite(v[u])) return LU_NONFINITE;  // entries before it are written, as the reference does
      a.x[(int64_t)qk[u] * B + c] = v[u];
    }
  }
  for (; k < n; k++) {
    T v = y[(int64_t)k * B + c];
    if (!lu_finite(v)) return LU_NONFINITE;
    a.x[(int64_t)q[k] * B + c] = v;
  }
  return LU_OK;
}

// ---- Mixed precision: FP32 factors, FP64 iterative refinement ----------------
//
// lu_factor_mixed factors the same permuted matrix as lu_factor with the same
// reach lists and the same pivot rule, but holds the working column, L, U and
// the pivots in single precision (FP32 for real systems, complex64 for AC
// systems: half the bytes of the factors). lu_solve_mixed solves with those
// factors and refines the solution with double-precision residuals
// r = b - A·y computed from the stamped FP64 / complex128 values, until the
// correction is at round-off. The double-precision path stays the reference:
// every case the mixed path cannot certify returns LU_UNSURE and is redone by
// lu_factor + lu_solve.
//
// Certified pivot sequence. Every single-precision column entry x_r carries a
// first-order bound e_r on its distance from the exact value of the same
// operations: conversion of A(r,k) (u·|a|, u = 2^-24), and for each update
// x_i -= l·x_j the propagated |l|·e_j + δl·|x_j| plus the rounding of the
// product and the difference, where δl is the stored error bound of the
// single-precision L entry (it follows from the bounds of its numerator and of
// the pivot at the column that produced it); complex products, quotients and
// magnitudes carry their larger rounding constants. The double-precision
// factorization's values lie well inside these bounds (its own rounding is
// 2^29 times smaller), so with intervals |x_r| ± kMixSafety·e_r a column's
// pivot is accepted only when the reference rule evaluated anywhere inside
// the intervals picks the same row: the argmax must win by more than both
// intervals, the diagonal preference (|x_k| >= 1e-3·best,
// solver/sparse.go:274-277) must hold or fail for every value in them, and
// the largest candidate must clear the singularity threshold
// tiny = colMax·n·1e-14, computed from the double-precision values as the
// reference does. The pivot rows of an accepted mixed factorization therefore
// equal the double-precision factorization's, which the parity tests check
// row for row.
//
// Storage inside the double-precision workspace: each single-precision value
// lives in the first half of its own entry's double-precision slot (entry i of
// lane c at i*B + c, as everywhere else), and a value's error bound in the
// slot's second half, so no lane's single-precision data ever overlaps
// another lane's entries: the column values and bounds in xw, the L entries
// and bounds in lx, U in ux, the pivots in the first halves of udiag. The
// refinement keeps w (permuted right-hand side), xw (the solution y) and x
// (the residual, until the solution is written) in double precision; the
// single-precision right-hand side of each pass is the first half of x's
// slots (formed from the residual entry by entry) and the single-precision
// correction the second half of udiag's.
constexpr float kMixU = 5.9604645e-8f;  // 2^-24, single-precision unit round-off
constexpr float kMixSafety = 4.0f;      // margin on the first-order error bounds
constexpr int kMixMaxIter = 8;

struct fcplx {
  float re, im;
};
KH_HD inline fcplx operator*(fcplx a, fcplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
KH_HD inline fcplx operator-(fcplx a, fcplx b) { return {a.re - b.re, a.im - b.im}; }

// Single-precision partner of each value type, with the rounding constants
// (in units of u) of the operations the mixed factorization uses.
template <typename T> struct Mix;
template <> struct Mix<double> {
  using Lo = float;
  static constexpr float kMul = 1.0f, kDiv = 1.0f, kMag = 0.0f;
  KH_HD static float lo(double v) { return (float)v; }
  KH_HD static double hi(float v) { return (double)v; }
  KH_HD static float conv_err(double v, float f) { return (float)fabs(v - (double)f); }
  KH_HD static float mag(float v) { return fabsf(v); }
  KH_HD static float mag_err(float, float) { return 0.0f; }
  KH_HD static float div(float a, float b) { return a / b; }
  KH_HD static bool zero(float v) { return v == 0.0f; }
  KH_HD static double nrm(double v) { return fabs(v); }
  KH_HD static bool in_range(double v) {
    const double av = fabs(v);
    return av == 0 || (av <= 1e30 && av >= 1e-30);
  }
};
template <> struct Mix<kcplx> {
  using Lo = fcplx;
  // complex product 2√2 u, Smith's quotient ~6u, scaled magnitude ~3u
  static constexpr float kMul = 3.0f, kDiv = 8.0f, kMag = 3.0f;
  KH_HD static fcplx lo(kcplx v) { return {(float)v.re, (float)v.im}; }
  KH_HD static kcplx hi(fcplx v) { return {(double)v.re, (double)v.im}; }
  KH_HD static float conv_err(kcplx v, fcplx f) {
    return (float)(fabs(v.re - (double)f.re) + fabs(v.im - (double)f.im));
  }
  // the scaled magnitude is exact for a purely real or imaginary value
  KH_HD static float mag_err(fcplx v, float av) {
    return (v.re != 0.0f && v.im != 0.0f) ? kMag * kMixU * av : 0.0f;
  }
  KH_HD static float mag(fcplx v) {
    float p = fabsf(v.re), q = fabsf(v.im);
    if (p < q) {
      const float t = p;
      p = q;
      q = t;
    }
    if (p == 0.0f) return 0.0f;
    q = q / p;
    return p * sqrtf(1.0f + q * q);
  }
  KH_HD static fcplx div(fcplx n, fcplx m) {
    if (fabsf(m.re) >= fabsf(m.im)) {
      const float ratio = m.im / m.re, denom = m.re + ratio * m.im;
      return {(n.re + n.im * ratio) / denom, (n.im - n.re * ratio) / denom};
    }
    const float ratio = m.re / m.im, denom = m.im + ratio * m.re;
    return {(n.re * ratio + n.im) / denom, (n.im * ratio - n.re) / denom};
  }
  KH_HD static bool zero(fcplx v) { return v.re == 0.0f && v.im == 0.0f; }
  KH_HD static double nrm(kcplx v) { return fmax(fabs(v.re), fabs(v.im)); }
  KH_HD static bool in_range(kcplx v) { return Mix<double>::in_range(v.re) && Mix<double>::in_range(v.im); }
};

// The single-precision value held in the first half of slot p of a T array,
// and the float held in its second half.
template <typename T>
KH_HD inline typename Mix<T>::Lo& mix_lo(T* base, int64_t p) {
  return *reinterpret_cast<typename Mix<T>::Lo*>(base + p);
}
template <typename T>
KH_HD inline float& mix_err(T* base, int64_t p) {
  return *reinterpret_cast<float*>(reinterpret_cast<char*>(base + p) + sizeof(typename Mix<T>::Lo));
}
template <typename T>
KH_HD inline typename Mix<T>::Lo& mix_hi_lo(T* base, int64_t p) {  // a Lo in the slot's second half
  return *reinterpret_cast<typename Mix<T>::Lo*>(reinterpret_cast<char*>(base + p) + sizeof(typename Mix<T>::Lo));
}

template <typename T, bool Unit = false>
KH_HD inline int lu_factor_mixed(const LuArgsT<T>& a, int64_t c) {
  using M = Mix<T>;
  using Lo = typename M::Lo;
  const int64_t B = Unit ? 1 : a.B;
  if (Unit) c = 0;
  const int n = a.n;
  const int32_t pidc = a.pid[c];
  const PatDesc pd = a.pats[pidc];
  const int32_t* ap = a.pool + pd.ap;
  const int32_t* arow = a.pool + pd.arow;
  const int32_t* aslot = a.pool + pd.aslot;
  T* const xw = a.xw;
  T* const lx = a.lx;
  T* const ux = a.ux;
  T* const ud = a.udiag;
  double scale = 0.0;
  for (int s = 0; s < a.nslots; s++) {
    const T v = a.val[(int64_t)s * B + c];
    if (!lu_finite(v)) return LU_UNSURE;  // the double-precision path reports ErrNonFinite
    const double av = lu_mag(v);
    if (av > scale) scale = av;
    if (!M::in_range(v)) return LU_UNSURE;  // outside single precision's comfortable range
  }
  if (scale == 0) return LU_UNSURE;  // the double-precision path reports ErrSingular
  bool reuse = a.cpid[c] == pidc;
  a.cpid[c] = -1;
  for (int i = 0; i < n; i++) a.pinv[(int64_t)i * B + c] = -1;
  if (!reuse)
    for (int i = 0; i < n; i++) a.mark[(int64_t)i * B + c] = 0;
  int64_t nl = 0, nu = 0, nr = 0;
  for (int k = 0; k < n; k++) {
    a.lp[(int64_t)k * B + c] = (int32_t)nl;
    a.up[(int64_t)k * B + c] = (int32_t)nu;
    int64_t r0 = nr, r1;
    bool diag_in_reach;
    if (reuse) {
      r1 = a.rp[(int64_t)(k + 1) * B + c];
      diag_in_reach = a.dk[(int64_t)k * B + c] != 0;
    } else {
      int top = lu_reach<T, Unit>(a, c, k, ap, arow);
      r1 = r0 + (n - top);
      if (r1 > a.capR) return LU_CAPACITY;
      for (int p = top; p < n; p++) a.rch[(r0 + p - top) * B + c] = a.xi[(int64_t)p * B + c];
      diag_in_reach = a.mark[(int64_t)k * B + c] == k + 1;
      a.rp[(int64_t)(k + 1) * B + c] = (int32_t)r1;
      a.dk[(int64_t)k * B + c] = diag_in_reach ? 1 : 0;
    }
    nr = r1;
    for (int64_t p = r0; p < r1; p++) {
      const int64_t ir = (int64_t)a.rch[p * B + c] * B + c;
      mix_lo(xw, ir) = M::lo(lu_zero_value<T>());
      mix_err(xw, ir) = 0.0f;
    }
    double colMax = 0.0;
    for (int p = ap[k]; p < ap[k + 1]; p++) {
      const T v = a.val[(int64_t)aslot[p] * B + c];
      const int64_t ir = (int64_t)arow[p] * B + c;
      const Lo vf = M::lo(v);
      mix_lo(xw, ir) = vf;
      mix_err(xw, ir) = M::conv_err(v, vf);
      const double av = lu_mag(v);
      if (av > colMax) colMax = av;
    }
    const double tiny = colMax * (double)n * 1e-14;
    for (int64_t p = r0; p < r1; p++) {
      const int32_t r = a.rch[p * B + c];
      const int32_t j = a.pinv[(int64_t)r * B + c];
      if (j < 0) continue;
      const Lo xj = mix_lo(xw, (int64_t)r * B + c);
      const float ej = mix_err(xw, (int64_t)r * B + c);
      if (M::zero(xj) && ej == 0.0f) continue;
      const float axj = M::mag(xj);
      const int32_t q1 = a.lp[(int64_t)(j + 1) * B + c];
      for (int32_t q = a.lp[(int64_t)j * B + c]; q < q1; q++) {
        const int64_t ix = (int64_t)a.li[(int64_t)q * B + c] * B + c;
        const Lo l = mix_lo(lx, (int64_t)q * B + c);
        const float al = M::mag(l);
        const Lo prod = l * xj;
        const Lo xn = mix_lo(xw, ix) - prod;
        mix_err(xw, ix) = mix_err(xw, ix) + al * ej + mix_err(lx, (int64_t)q * B + c) * axj + kMixU * (M::kMul * al * axj + M::mag(xn));
        mix_lo(xw, ix) = xn;
      }
    }
    int32_t piv = -1, rhi = -1;
    float best = -1.0f, maxhi = -1.0f, maxhi2 = -1.0f, maxlo = -1.0f;
    for (int64_t p = r0; p < r1; p++) {
      const int32_t r = a.rch[p * B + c];
      const int64_t ir = (int64_t)r * B + c;
      const int32_t pr = a.pinv[ir];
      if (pr >= 0) {
        if (nu >= a.capU) return LU_CAPACITY;
        a.ui[nu * B + c] = pr;
        mix_lo(ux, nu * B + c) = mix_lo(xw, ir);
        nu++;
        continue;
      }
      const float av = M::mag(mix_lo(xw, ir)), err = kMixSafety * (mix_err(xw, ir) + M::mag_err(mix_lo(xw, ir), av));
      if (av > best || (av == best && r < piv)) {
        best = av;
        piv = r;
      }
      const float hi = av + err, lo = av - err;
      if (hi > maxhi) {
        maxhi2 = maxhi;
        maxhi = hi;
        rhi = r;
      } else if (hi > maxhi2) {
        maxhi2 = hi;
      }
      if (lo > maxlo) maxlo = lo;
    }
    if (piv < 0 || !((double)maxlo > tiny)) return LU_UNSURE;
    bool diag = false;
    if (a.pinv[(int64_t)k * B + c] < 0 && diag_in_reach) {
      const float ak = M::mag(mix_lo(xw, (int64_t)k * B + c));
      const float ek = kMixSafety * (mix_err(xw, (int64_t)k * B + c) + M::mag_err(mix_lo(xw, (int64_t)k * B + c), ak));
      const float lok = ak - ek, hik = ak + ek;
      if ((double)lok >= 1e-3 * (double)maxhi && (double)lok > tiny) {
        diag = true;
      } else if (!((double)hik < 1e-3 * (double)maxlo || (double)hik <= tiny)) {
        return LU_UNSURE;
      }
    }
    if (diag) {
      piv = k;
    } else {
      // The argmax: every other candidate's interval lies below the winner's,
      // or both values are exact and equal and the winner has the lower row
      // (the reference's tie-break, solver/sparse.go:262-268).
      const float wp = kMixSafety * (mix_err(xw, (int64_t)piv * B + c) + M::mag_err(mix_lo(xw, (int64_t)piv * B + c), best));
      const float lop = best - wp;
      if (!(lop > (rhi == piv ? maxhi2 : maxhi))) {
        for (int64_t p = r0; p < r1; p++) {
          const int32_t r = a.rch[p * B + c];
          const int64_t ir = (int64_t)r * B + c;
          if (r == piv || a.pinv[ir] >= 0) continue;
          const float av = M::mag(mix_lo(xw, ir)), wr = kMixSafety * (mix_err(xw, ir) + M::mag_err(mix_lo(xw, ir), av));
          if (lop > av + wr) continue;
          if (wp == 0.0f && wr == 0.0f && av == best && piv < r) continue;
          return LU_UNSURE;
        }
      }
    }
    if (reuse && piv != a.prow[(int64_t)k * B + c]) {
      reuse = false;
      for (int i = 0; i < n; i++) a.mark[(int64_t)i * B + c] = 0;
    }
    const Lo pv = mix_lo(xw, (int64_t)piv * B + c);
    const float apv = M::mag(pv), epv = mix_err(xw, (int64_t)piv * B + c);
    mix_lo(ud, (int64_t)k * B + c) = pv;
    a.pinv[(int64_t)piv * B + c] = k;
    a.prow[(int64_t)k * B + c] = piv;
    for (int64_t p = r0; p < r1; p++) {
      const int32_t r = a.rch[p * B + c];
      const int64_t ir = (int64_t)r * B + c;
      if (a.pinv[ir] < 0) {
        if (nl >= a.capL) return LU_CAPACITY;
        const Lo l = M::div(mix_lo(xw, ir), pv);
        const float al = M::mag(l);
        a.li[nl * B + c] = r;
        mix_lo(lx, nl * B + c) = l;
        mix_err(lx, nl * B + c) = (mix_err(xw, ir) + al * epv) / apv + M::kDiv * kMixU * al;
        nl++;
      }
    }
  }
  a.lp[(int64_t)n * B + c] = (int32_t)nl;
  a.up[(int64_t)n * B + c] = (int32_t)nu;
  a.rp[c] = 0;
  a.cpid[c] = pidc;
  return LU_OK;
}

// Solve with the single-precision factors of lu_factor_mixed and refine in
// double precision. Each pass forms the residual of the current solution y
// against the stamped double-precision values (the first pass starts from
// y = 0, r = b), rounds it to single precision, solves for the correction with
// the single-precision factors and adds it to y. The corrections shrink
// geometrically; with rho the ratio of the last two, what remains after a
// pass is about rho times its correction. It stops when that is within 2^-46
// (1.4e-14) of every component plus 2^-60 of the largest component (for
// components at zero), or when the corrections have reached the residual's
// own round-off (no longer halving, below 1e-12 of the solution). Divergence,
// a non-finite value or kMixMaxIter passes return LU_UNSURE. *iters receives
// the number of single-precision solves.
template <typename T, bool Unit = false>
KH_HD inline int lu_solve_mixed(const LuArgsT<T>& a, int64_t c, int* iters) {
  using M = Mix<T>;
  using Lo = typename M::Lo;
  const int64_t B = Unit ? 1 : a.B;
  if (Unit) c = 0;
  const int n = a.n;
  const PatDesc pd = a.pats[a.pid[c]];
  const int32_t* q = a.pool + pd.q;
  const int32_t* qinv = a.pool + pd.qinv;
  const int32_t* ap = a.pool + pd.ap;
  const int32_t* arow = a.pool + pd.arow;
  const int32_t* aslot = a.pool + pd.aslot;
  T* const lx = a.lx;
  T* const ux = a.ux;
  T* const ud = a.udiag;
  T* w = a.w;
  T* y = a.xw;
  T* r = a.x;
  for (int i = 0; i < n; i++) w[(int64_t)qinv[i] * B + c] = a.rhs[(int64_t)i * B + c];
  double prev = __builtin_huge_val();
  int it = 0;
  for (;;) {
    if (it == 0) {
      for (int i = 0; i < n; i++) mix_lo(r, (int64_t)i * B + c) = M::lo(w[(int64_t)i * B + c]);
    } else {
      for (int i = 0; i < n; i++) r[(int64_t)i * B + c] = w[(int64_t)i * B + c];
      for (int k = 0; k < n; k++) {
        const T yk = y[(int64_t)k * B + c];
        if (lu_zero(yk)) continue;
        for (int p = ap[k]; p < ap[k + 1]; p++) {
          const int64_t ir = (int64_t)arow[p] * B + c;
          const T prod = a.val[(int64_t)aslot[p] * B + c] * yk;
          r[ir] = r[ir] - prod;
        }
      }
      for (int i = 0; i < n; i++) mix_lo(r, (int64_t)i * B + c) = M::lo(r[(int64_t)i * B + c]);
    }
    for (int k = 0; k < n; k++) {
      const Lo zk = mix_lo(r, (int64_t)a.prow[(int64_t)k * B + c] * B + c);
      mix_hi_lo(ud, (int64_t)k * B + c) = zk;
      if (M::zero(zk)) continue;
      const int32_t p1 = a.lp[(int64_t)(k + 1) * B + c];
      for (int32_t p = a.lp[(int64_t)k * B + c]; p < p1; p++) {
        const int64_t ix = (int64_t)a.li[(int64_t)p * B + c] * B + c;
        mix_lo(r, ix) = mix_lo(r, ix) - mix_lo(lx, (int64_t)p * B + c) * zk;
      }
    }
    for (int k = n - 1; k >= 0; k--) {
      const Lo zk = M::div(mix_hi_lo(ud, (int64_t)k * B + c), mix_lo(ud, (int64_t)k * B + c));
      mix_hi_lo(ud, (int64_t)k * B + c) = zk;
      if (M::zero(zk)) continue;
      const int32_t p1 = a.up[(int64_t)(k + 1) * B + c];
      for (int32_t p = a.up[(int64_t)k * B + c]; p < p1; p++) {
        const int64_t ix = (int64_t)a.ui[(int64_t)p * B + c] * B + c;
        mix_hi_lo(ud, ix) = mix_hi_lo(ud, ix) - mix_lo(ux, (int64_t)p * B + c) * zk;
      }
    }
    double ymax = 0.0, dmax = 0.0;
    for (int k = 0; k < n; k++) {
      const T dz = M::hi(mix_hi_lo(ud, (int64_t)k * B + c));
      const T yk = it == 0 ? dz : y[(int64_t)k * B + c] + dz;
      if (!lu_finite(yk)) return LU_UNSURE;
      y[(int64_t)k * B + c] = yk;
      ymax = fmax(ymax, M::nrm(yk));
      dmax = fmax(dmax, M::nrm(dz));
    }
    it++;
    const double rel = ymax > 0 ? dmax / ymax : 0.0;
    if (it >= 2) {
      // rho: contraction of the corrections; what remains after this pass is
      // about rho times this pass's correction.
      const double rho = fmin(1.0, 2.0 * rel / prev);
      bool tight = true;
      for (int k = 0; k < n && tight; k++) {
        const double yk = M::nrm(y[(int64_t)k * B + c]);
        const double dz = M::nrm(M::hi(mix_hi_lo(ud, (int64_t)k * B + c)));
        tight = rho * dz <= 0x1p-46 * yk + 0x1p-60 * ymax;
      }
      if (tight) break;
      if (rel <= 1e-12 && rel > 0.5 * prev) break;  // at the residual's round-off
      if (rel > prev) return LU_UNSURE;             // diverging
    }
    prev = it == 1 ? 1.0 : rel;
    if (it >= kMixMaxIter) return LU_UNSURE;
  }
  for (int k = 0; k < n; k++) a.x[(int64_t)q[k] * B + c] = y[(int64_t)k * B + c];
  *iters = it;
  return LU_OK;
}

// Dense backend: Dense.Solve / CDense.Factor + SolveFactored
// (solver/dense.go:80-178, solver/cdense.go:92-193), which Options.Backend
// Auto picks below 30 unknowns (solver/matrix.go:77-85). The lane's slot
// values are scattered into a row-major matrix (lx), which is factored in
// place by Doolittle elimination with partial pivoting: the largest magnitude
// at or below the diagonal, the first one on ties; a pivot at most n·1e-14 of
// its column's largest original entry is singular. Forward substitution with
// the unit lower factor, back substitution with the upper one; the solution is
// written to x before its finiteness is checked, as Dense.Solve does.
// Workspace: lx [n*n], udiag (column maxima), w (y), prow (pivot rows).
template <typename T, bool Unit = false>
KH_HD inline int lu_dense(const LuArgsT<T>& a, int64_t c, const int32_t* slot_i, const int32_t* slot_j) {
  const int64_t B = Unit ? 1 : a.B;
  if (Unit) c = 0;
  const int n = a.n;
  T* m = a.lx;
  auto M = [&](int i, int j) -> T& { return m[((int64_t)i * n + j) * B + c]; };
  for (int64_t k = 0; k < (int64_t)n * n; k++) m[k * B + c] = lu_zero_value<T>();
  for (int s = 0; s < a.nslots; s++) M(slot_i[s], slot_j[s]) = a.val[(int64_t)s * B + c];
  double scale = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const T v = M(i, j);
      if (!lu_finite(v)) return LU_NONFINITE;
      const double mv = lu_mag(v);
      if (mv > scale) scale = mv;
    }
  if (scale == 0) {
    a.err_step[c] = 0;
    return LU_SINGULAR;
  }
  // Column maxima of the original matrix, in the real parts of udiag.
  auto COL = [&](int j) -> double& { return lu_re(a.udiag[(int64_t)j * B + c]); };
  for (int j = 0; j < n; j++) a.udiag[(int64_t)j * B + c] = lu_zero_value<T>();
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double mv = lu_mag(M(i, j));
      if (mv > COL(j)) COL(j) = mv;
    }
  for (int k = 0; k < n; k++) {
    int p = k;
    double maxv = lu_mag(M(k, k));
    for (int i = k + 1; i < n; i++) {
      const double v = lu_mag(M(i, k));
      if (v > maxv) {
        maxv = v;
        p = i;
      }
    }
    if (maxv <= COL(k) * (double)n * 1e-14) {
      a.err_step[c] = k;
      return LU_SINGULAR;
    }
    a.prow[(int64_t)k * B + c] = p;
    if (p != k)
      for (int j = 0; j < n; j++) {
        const T t = M(k, j);
        M(k, j) = M(p, j);
        M(p, j) = t;
      }
    const T pivot = M(k, k);
    for (int i = k + 1; i < n; i++) {
      const T f = lu_div(M(i, k), pivot);
      M(i, k) = f;
      if (lu_zero(f)) continue;
      for (int j = k + 1; j < n; j++) {
        const T prod = f * M(k, j);
        M(i, j) = M(i, j) - prod;
      }
    }
  }
  T* y = a.w;
  for (int i = 0; i < n; i++) y[(int64_t)i * B + c] = a.rhs[(int64_t)i * B + c];
  for (int k = 0; k < n; k++) {
    const int p = a.prow[(int64_t)k * B + c];
    if (p != k) {
      const T t = y[(int64_t)k * B + c];
      y[(int64_t)k * B + c] = y[(int64_t)p * B + c];
      y[(int64_t)p * B + c] = t;
    }
  }
  for (int i = 0; i < n; i++) {
    T sum = y[(int64_t)i * B + c];
    for (int j = 0; j < i; j++) {
      const T prod = M(i, j) * y[(int64_t)j * B + c];
      sum = sum - prod;
    }
    y[(int64_t)i * B + c] = sum;
  }
  for (int i = n - 1; i >= 0; i--) {
    T sum = y[(int64_t)i * B + c];
    for (int j = i + 1; j < n; j++) {
      const T prod = M(i, j) * a.x[(int64_t)j * B + c];
      sum = sum - prod;
    }
    a.x[(int64_t)i * B + c] = lu_div(sum, M(i, i));
  }
  for (int i = 0; i < n; i++)
    if (!lu_finite(a.x[(int64_t)i * B + c])) return LU_NONFINITE;
  return LU_OK;
}

}  // namespace kh
