// kh_tear.h — torn solves of MNA Jacobians on the GPU: kron.JacobianTearing
// (kron/jacobian.go:66-712) as a batched solver, one lane per system.
//
// A torn matrix's unknowns fall into pieces: the connected components of its
// pattern once the cut node pairs' couplings are removed. With B the
// block-diagonal part and the coupling entries E = P·M·Q (P selecting their
// rows, Q their columns),
//   z = B⁻¹·b                       each piece on its own
//   (I + M·Q·B⁻¹·P)·w = M·Q·z       interconnection, order = coupling rows
//   x = z − B⁻¹·P·w
// (Aitchison's algorithm). Each piece and the interconnection are factored
// by an equilibrated dense LU with partial pivoting; a piece whose block is
// bit-identical to the one it last factored keeps its factors (or, with a
// reuse tolerance, one that moved by at most that fraction of its largest
// entry); every solution is refined against the whole matrix to a
// componentwise backward error of 1e-14 and accepted at 1e-12. A system that
// is refused (a singular piece, no convergence) is marked for the untorn LU,
// which the caller runs next (solver.Auto: dense below 30 unknowns, else the
// sparse Gilbert–Peierls LU).
//
// The plan (pieces, ports, coupling slots) depends only on the pattern and
// the cut and is shared by every system of a batch; it is built on the host
// (tear_plan). Each system keeps its pieces' factors, last block values and
// port columns between calls in three state arrays laid out
// structure-of-arrays over the systems (word e of system c at [e*B + c]), so
// the lanes of a wavefront touch consecutive words.
#pragma once
#include <algorithm>
#include <cstdint>
#include <unordered_map>
#include <utility>
#include <vector>

#include "kh_lu.h"

namespace kh {

constexpr int kTearMaxRefine = 12;              // maxRefine, kron/diakoptics.go:83
constexpr double kTearRefineDone = 1e-14;       // refineDone, :84
constexpr double kTearRefineAccept = 1e-12;     // refineAccept, :85
constexpr int kTearMaxPiece = 4096;             // solver.MaxDense (solver/dense.go:12): a larger piece is refused
constexpr int TEAR_FALLBACK = 17;               // the system needs the untorn LU

// Counters one lane adds to, per system ([kTearCounters][B] int64).
enum { TC_SOLVES = 0, TC_FACTORS, TC_REUSES, TC_STALE, TC_FALLBACKS, TC_REFINES, kTearCounters };

// One piece: m unknowns idx[i0 .. i0+m), ns block slots (slot ids, local row
// and column) at s0, np ports (indices into the coupling rows) at p0, and
// the offsets of its state: T words lu (m*m), last (ns), W (np*m); double
// words rs, cs (m each); int words piv (m), have (1).
struct TearPiece {
  int32_t m, i0, ns, s0, np, p0;
  int64_t lu, last, W, rs, cs, piv, have;
};

// The plan as the kernels see it (device pointers).
struct TearPlanView {
  int32_t n, nslots, npieces, icN, noff, maxm;
  int64_t nT, nD, nI;              // state words per system (T, double, int)
  int64_t ic_lu, ic_rs, ic_cs, ic_piv;  // interconnection factors
  int64_t v_r, v_dx, v_h, v_w, v_b, v_xl;  // T work vectors: r, dx [n]; h, w [icN]; b, xl [maxm]
  int64_t v_den, v_col;            // double work vectors: den [n], col [max(maxm, icN)]
  const int32_t* ri;       // [nslots] row of each slot
  const int32_t* ci;       // [nslots] column
  const int32_t* pieceOf;  // [n]
  const int32_t* local;    // [n] index within its piece
  const int32_t* rowIdx;   // [n] coupling-row index, or -1
  const int32_t* off;      // [noff] coupling slots
  const int32_t* rows;     // [icN] global row of each coupling row
  const int32_t* idx;      // pieces' unknowns, concatenated
  const int32_t* pslot;    // pieces' block slots, concatenated
  const int32_t* pli;      // their local rows
  const int32_t* plj;      // and local columns
  const int32_t* ports;    // pieces' ports, concatenated
  const TearPiece* pc;     // [npieces]
};

// math.Max: NaN if either argument is NaN, +Inf if either is +Inf, +0 over -0.
KH_HD inline double go_max(double a, double b) {
  if (isinf(a) && a > 0) return a;
  if (isinf(b) && b > 0) return b;
  if (isnan(a) || isnan(b)) return __builtin_nan("");
  if (a == 0 && a == b) return signbit(a) ? b : a;
  return a > b ? a : b;
}

template <typename T> KH_HD inline T tear_from(double v);
template <> KH_HD inline double tear_from<double>(double v) { return v; }
template <> KH_HD inline kcplx tear_from<kcplx>(double v) { return {v, 0.0}; }
KH_HD inline bool tear_ne(double a, double b) { return a != b; }
KH_HD inline bool tear_ne(kcplx a, kcplx b) { return a.re != b.re || a.im != b.im; }
// math.IsNaN(f) || math.IsInf(f, 0) of f = |v| (cmplx.Abs for complex values)
template <typename T> KH_HD inline bool tear_bad(T v) {
  const double f = lu_mag(v);
  return isnan(f) || isinf(f);
}

// One system's torn solve on one lane. val [nslots][B], rhs and x [n][B] are
// the system's values, right-hand side and solution; tv, dv, iv its state.
template <typename T>
struct TearLane {
  const TearPlanView& P;
  int64_t B, c;
  const T* val;
  const T* rhs;
  T* x;
  T* tv;
  double* dv;
  int32_t* iv;
  int64_t* cnt;
  double reuseTol;

  KH_HD T& TV(int64_t e) const { return tv[e * B + c]; }
  KH_HD double& DV(int64_t e) const { return dv[e * B + c]; }
  KH_HD int32_t& IV(int64_t e) const { return iv[e * B + c]; }
  KH_HD T V(int s) const { return val[(int64_t)s * B + c]; }
  KH_HD void count(int k, int64_t d = 1) const { cnt[(int64_t)k * B + c] += d; }

  // denseLU.factor (kron/jacobian.go:611-672) of an m×m matrix at lu with
  // row and column scales rs, cs and pivot rows piv; ent(f) calls f(i, j, v)
  // for every entry in order (duplicates accumulate). false: singular.
  template <class Ent>
  KH_HD bool dense_factor(int m, int64_t lu, int64_t rs, int64_t cs, int64_t piv, Ent ent) const {
    const T zero = lu_zero_value<T>();
    for (int64_t k = 0; k < (int64_t)m * m; k++) TV(lu + k) = zero;
    for (int i = 0; i < m; i++) {
      DV(rs + i) = 0;
      DV(cs + i) = 0;
      DV(P.v_col + i) = 0;
    }
    ent([&](int i, int j, T v) { DV(rs + i) = go_max(DV(rs + i), lu_mag(v)); });
    for (int i = 0; i < m; i++) {
      if (DV(rs + i) == 0) return false;  // equation i is empty
      DV(rs + i) = 1 / DV(rs + i);
    }
    ent([&](int i, int j, T v) { DV(cs + j) = go_max(DV(cs + j), lu_mag(v) * DV(rs + i)); });
    for (int j = 0; j < m; j++) {
      if (DV(cs + j) == 0) return false;  // unknown j appears in no equation
      DV(cs + j) = 1 / DV(cs + j);
    }
    ent([&](int i, int j, T v) {
      const T s = v * tear_from<T>(DV(rs + i)) * tear_from<T>(DV(cs + j));
      T& a = TV(lu + (int64_t)i * m + j);
      a = a + s;
    });
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) DV(P.v_col + j) = go_max(DV(P.v_col + j), lu_mag(TV(lu + (int64_t)i * m + j)));
    for (int k = 0; k < m; k++) {
      int p = k;
      double maxv = lu_mag(TV(lu + (int64_t)k * m + k));
      for (int i = k + 1; i < m; i++) {
        const double v = lu_mag(TV(lu + (int64_t)i * m + k));
        if (v > maxv) {
          maxv = v;
          p = i;
        }
      }
      if (maxv <= DV(P.v_col + k) * (double)m * 1e-14) return false;
      IV(piv + k) = p;
      if (p != k)
        for (int j = 0; j < m; j++) {
          const T t = TV(lu + (int64_t)k * m + j);
          TV(lu + (int64_t)k * m + j) = TV(lu + (int64_t)p * m + j);
          TV(lu + (int64_t)p * m + j) = t;
        }
      const T pivot = TV(lu + (int64_t)k * m + k);
      for (int i = k + 1; i < m; i++) {
        const T f = lu_div(TV(lu + (int64_t)i * m + k), pivot);
        TV(lu + (int64_t)i * m + k) = f;
        if (lu_zero(f)) continue;
        for (int j = k + 1; j < m; j++) {
          const T prod = f * TV(lu + (int64_t)k * m + j);
          T& a = TV(lu + (int64_t)i * m + j);
          a = a - prod;
        }
      }
    }
    return true;
  }

  // denseLU.solve (:676-712): x = A⁻¹·b for state vectors at b and xo (distinct).
  KH_HD void dense_solve(int m, int64_t lu, int64_t rs, int64_t cs, int64_t piv, int64_t b, int64_t xo) const {
    for (int i = 0; i < m; i++) TV(xo + i) = TV(b + i) * tear_from<T>(DV(rs + i));
    for (int k = 0; k < m; k++) {
      const int p = IV(piv + k);
      if (p != k) {
        const T t = TV(xo + k);
        TV(xo + k) = TV(xo + p);
        TV(xo + p) = t;
      }
    }
    for (int i = 0; i < m; i++) {
      T s = TV(xo + i);
      for (int j = 0; j < i; j++) {
        const T prod = TV(lu + (int64_t)i * m + j) * TV(xo + j);
        s = s - prod;
      }
      TV(xo + i) = s;
    }
    for (int i = m - 1; i >= 0; i--) {
      T s = TV(xo + i);
      for (int j = i + 1; j < m; j++) {
        const T prod = TV(lu + (int64_t)i * m + j) * TV(xo + j);
        s = s - prod;
      }
      TV(xo + i) = lu_div(s, TV(lu + (int64_t)i * m + i));
    }
    for (int j = 0; j < m; j++) TV(xo + j) = TV(xo + j) * tear_from<T>(DV(cs + j));
  }

  // apply (the closure in torn, :503-544): out = J⁻¹·in through the pieces
  // and the interconnection; in(g) reads, out(g) references unknown g.
  template <class In, class Out>
  KH_HD void apply(In in, Out out) const {
    const T zero = lu_zero_value<T>();
    for (int q = 0; q < P.npieces; q++) {
      const TearPiece pc = P.pc[q];
      bool nz = false;
      for (int li = 0; li < pc.m; li++) {
        const T v = in(P.idx[pc.i0 + li]);
        TV(P.v_b + li) = v;
        nz = nz || !lu_zero(v);
      }
      if (!nz) {
        for (int li = 0; li < pc.m; li++) out(P.idx[pc.i0 + li]) = zero;
        continue;
      }
      dense_solve(pc.m, pc.lu, pc.rs, pc.cs, pc.piv, P.v_b, P.v_xl);
      for (int li = 0; li < pc.m; li++) out(P.idx[pc.i0 + li]) = TV(P.v_xl + li);
    }
    if (P.icN == 0) return;
    for (int a = 0; a < P.icN; a++) TV(P.v_h + a) = zero;
    for (int k = 0; k < P.noff; k++) {
      const int s = P.off[k];
      const T prod = V(s) * out(P.ci[s]);
      T& h = TV(P.v_h + P.rowIdx[P.ri[s]]);
      h = h + prod;
    }
    dense_solve(P.icN, P.ic_lu, P.ic_rs, P.ic_cs, P.ic_piv, P.v_h, P.v_w);
    for (int q = 0; q < P.npieces; q++) {
      const TearPiece pc = P.pc[q];
      for (int k = 0; k < pc.np; k++) {
        const T wk = TV(P.v_w + P.ports[pc.p0 + k]);
        if (lu_zero(wk)) continue;
        for (int li = 0; li < pc.m; li++) {
          const T prod = wk * TV(pc.W + (int64_t)k * pc.m + li);
          T& z = out(P.idx[pc.i0 + li]);
          z = z - prod;
        }
      }
    }
  }

  // torn (:436-589): true when the torn solve succeeded (x then holds the
  // solution); *stale tells whether it used factors within the reuse tolerance.
  KH_HD bool torn(bool allowStale, bool* stale) const {
    const T zero = lu_zero_value<T>();
    *stale = false;
    for (int q = 0; q < P.npieces; q++) {
      const TearPiece pc = P.pc[q];
      const bool have = IV(pc.have) != 0;
      bool same = have;
      const bool small = have && allowStale && reuseTol > 0;
      double mx = 0, dmax = 0;
      for (int k = 0; k < pc.ns; k++) {
        const T v = V(P.pslot[pc.s0 + k]);
        const T last = TV(pc.last + k);
        if (tear_ne(v, last)) {
          same = false;
          if (small) dmax = go_max(dmax, lu_mag(v - last));
        }
        if (small) mx = go_max(mx, lu_mag(v));
      }
      if (same) {
        count(TC_REUSES);
        continue;
      }
      if (small && dmax <= reuseTol * mx) {
        count(TC_STALE);
        *stale = true;
        continue;
      }
      for (int k = 0; k < pc.ns; k++) TV(pc.last + k) = V(P.pslot[pc.s0 + k]);
      IV(pc.have) = 0;
      const bool ok = dense_factor(pc.m, pc.lu, pc.rs, pc.cs, pc.piv, [&](auto&& f) {
        for (int k = 0; k < pc.ns; k++) f(P.pli[pc.s0 + k], P.plj[pc.s0 + k], V(P.pslot[pc.s0 + k]));
      });
      if (!ok) return false;
      count(TC_FACTORS);
      for (int k = 0; k < pc.np; k++) {
        for (int li = 0; li < pc.m; li++) TV(P.v_b + li) = zero;
        TV(P.v_b + P.local[P.rows[P.ports[pc.p0 + k]]]) = tear_from<T>(1.0);
        dense_solve(pc.m, pc.lu, pc.rs, pc.cs, pc.piv, P.v_b, pc.W + (int64_t)k * pc.m);
      }
      IV(pc.have) = 1;
    }
    if (P.icN > 0) {
      // I + M·Q·B⁻¹·P: the identity first, then each coupling entry's
      // products with the port columns of its column's piece.
      const bool ok = dense_factor(P.icN, P.ic_lu, P.ic_rs, P.ic_cs, P.ic_piv, [&](auto&& f) {
        for (int a = 0; a < P.icN; a++) f(a, a, tear_from<T>(1.0));
        for (int k = 0; k < P.noff; k++) {
          const int s = P.off[k];
          const T v = V(s);
          if (lu_zero(v)) continue;
          const int col = P.ci[s];
          const TearPiece pc = P.pc[P.pieceOf[col]];
          const int lc = P.local[col];
          for (int t = 0; t < pc.np; t++) {
            const T g = TV(pc.W + (int64_t)t * pc.m + lc);
            if (!lu_zero(g)) f(P.rowIdx[P.ri[s]], P.ports[pc.p0 + t], v * g);
          }
        }
      });
      if (!ok) return false;
    }
    const int n = P.n;
    auto X = [&](int g) -> T& { return x[(int64_t)g * B + c]; };
    auto Bv = [&](int g) -> T { return rhs[(int64_t)g * B + c]; };
    apply(Bv, X);
    double berr = __builtin_huge_val();
    for (int pass = 0; pass <= kTearMaxRefine; pass++) {
      for (int i = 0; i < n; i++) {
        TV(P.v_r + i) = Bv(i);
        DV(P.v_den + i) = lu_mag(Bv(i));
      }
      for (int s = 0; s < P.nslots; s++) {
        const int i = P.ri[s];
        const T t = V(s) * X(P.ci[s]);
        TV(P.v_r + i) = TV(P.v_r + i) - t;
        DV(P.v_den + i) = DV(P.v_den + i) + lu_mag(t);
      }
      berr = 0;
      for (int i = 0; i < n; i++) {
        if (DV(P.v_den + i) > 0) berr = go_max(berr, lu_mag(TV(P.v_r + i)) / DV(P.v_den + i));
        else if (!lu_zero(TV(P.v_r + i))) berr = __builtin_huge_val();
      }
      if (berr <= kTearRefineDone || pass == kTearMaxRefine) break;
      apply([&](int g) -> T { return TV(P.v_r + g); }, [&](int g) -> T& { return TV(P.v_dx + g); });
      for (int i = 0; i < n; i++) X(i) = X(i) + TV(P.v_dx + i);
      count(TC_REFINES);
    }
    if (!(berr <= kTearRefineAccept)) return false;
    for (int i = 0; i < n; i++)
      if (tear_bad(X(i))) return false;
    return true;
  }

  // Solve (:393-416) up to the untorn fallback: LU_OK, LU_NONFINITE (a
  // non-finite value in the system), or TEAR_FALLBACK (torn solve refused:
  // the caller solves the system untorn). refused: the plan itself was
  // refused (a piece larger than solver.MaxDense).
  KH_HD int solve(bool refused) const {
    count(TC_SOLVES);
    for (int s = 0; s < P.nslots; s++)
      if (tear_bad(V(s))) return LU_NONFINITE;
    if (refused) {
      count(TC_FALLBACKS);
      return TEAR_FALLBACK;
    }
    bool stale = false;
    bool ok = torn(true, &stale);
    if (!ok && stale) {
      bool again = false;
      ok = torn(false, &again);
    }
    if (!ok) {
      count(TC_FALLBACKS);
      return TEAR_FALLBACK;
    }
    return LU_OK;
  }
};

// ---- host: the plan ------------------------------------------------------------------

// The plan of one pattern under one cut (buildPlan, kron/jacobian.go:298-389)
// and the state offsets of one system (kh_tear.h).
struct TearHost {
  int n = 0;
  int64_t nslots = 0;
  std::vector<int32_t> ri, ci, pieceOf, local, rowIdx, off, rows, idx, pslot, pli, plj, ports;
  std::vector<TearPiece> pcs;
  int icN = 0, maxm = 0;
  bool refused = false;  // a piece larger than solver.MaxDense: every solve is untorn
  TearPlanView v = {};   // offsets and sizes; the pointers are set on the device copy
};

inline TearHost tear_plan(int n, const std::vector<int32_t>& ri, const std::vector<int32_t>& ci,
                   const std::vector<std::pair<int32_t, int32_t>>& cuts) {
  TearHost h;
  h.n = n;
  h.nslots = (int64_t)ri.size();
  h.ri = ri;
  h.ci = ci;
  std::unordered_map<int64_t, bool> cut;
  for (const auto& c : cuts)
    if (c.first < n && c.second < n) {
      cut[(int64_t)c.first * n + c.second] = true;
      cut[(int64_t)c.second * n + c.first] = true;
    }
  // union-find over the uncut off-diagonal entries; the smaller root wins
  std::vector<int32_t> parent(n);
  for (int i = 0; i < n; i++) parent[i] = i;
  auto find = [&](int32_t i) {
    while (parent[i] != i) {
      parent[i] = parent[parent[i]];
      i = parent[i];
    }
    return i;
  };
  for (int64_t s = 0; s < h.nslots; s++) {
    const int32_t i = ri[s], j = ci[s];
    if (i == j || cut.count((int64_t)i * n + j)) continue;
    const int32_t a = find(i), b = find(j);
    if (a != b) {
      if (a < b) parent[b] = a;
      else parent[a] = b;
    }
  }
  // pieces in order of their first unknown; each piece's unknowns ascending
  h.pieceOf.assign(n, 0);
  h.local.assign(n, 0);
  h.rowIdx.assign(n, -1);
  std::vector<int32_t> root(n, -1);
  std::vector<std::vector<int32_t>> members;
  for (int32_t i = 0; i < n; i++) {
    const int32_t r = find(i);
    if (root[r] < 0) {
      root[r] = (int32_t)members.size();
      members.emplace_back();
    }
    h.pieceOf[i] = root[r];
    h.local[i] = (int32_t)members[root[r]].size();
    members[root[r]].push_back(i);
  }
  const int np = (int)members.size();
  std::vector<std::vector<int32_t>> bslot(np), bli(np), blj(np), bport(np);
  for (int64_t s = 0; s < h.nslots; s++) {
    const int32_t i = ri[s], j = ci[s];
    if (h.pieceOf[i] == h.pieceOf[j]) {
      const int q = h.pieceOf[i];
      bslot[q].push_back((int32_t)s);
      bli[q].push_back(h.local[i]);
      blj[q].push_back(h.local[j]);
      continue;
    }
    h.off.push_back((int32_t)s);
    if (h.rowIdx[i] < 0) {
      h.rowIdx[i] = (int32_t)h.rows.size();
      h.rows.push_back(i);
      bport[h.pieceOf[i]].push_back(h.rowIdx[i]);
    }
  }
  h.icN = (int)h.rows.size();
  int64_t nT = 0, nD = 0, nI = 0;
  for (int q = 0; q < np; q++) {
    TearPiece pc = {};
    pc.m = (int32_t)members[q].size();
    if (pc.m > kTearMaxPiece) h.refused = true;
    h.maxm = std::max(h.maxm, pc.m);
    pc.i0 = (int32_t)h.idx.size();
    h.idx.insert(h.idx.end(), members[q].begin(), members[q].end());
    pc.ns = (int32_t)bslot[q].size();
    pc.s0 = (int32_t)h.pslot.size();
    h.pslot.insert(h.pslot.end(), bslot[q].begin(), bslot[q].end());
    h.pli.insert(h.pli.end(), bli[q].begin(), bli[q].end());
    h.plj.insert(h.plj.end(), blj[q].begin(), blj[q].end());
    pc.np = (int32_t)bport[q].size();
    pc.p0 = (int32_t)h.ports.size();
    h.ports.insert(h.ports.end(), bport[q].begin(), bport[q].end());
    pc.lu = nT; nT += (int64_t)pc.m * pc.m;
    pc.last = nT; nT += pc.ns;
    pc.W = nT; nT += (int64_t)pc.np * pc.m;
    pc.rs = nD; nD += pc.m;
    pc.cs = nD; nD += pc.m;
    pc.piv = nI; nI += pc.m;
    pc.have = nI; nI += 1;
    h.pcs.push_back(pc);
  }
  TearPlanView& v = h.v;
  v.n = n;
  v.nslots = (int32_t)h.nslots;
  v.npieces = np;
  v.icN = h.icN;
  v.noff = (int32_t)h.off.size();
  v.maxm = h.maxm;
  v.ic_lu = nT; nT += (int64_t)h.icN * h.icN;
  v.v_r = nT; nT += n;
  v.v_dx = nT; nT += n;
  v.v_h = nT; nT += h.icN;
  v.v_w = nT; nT += h.icN;
  v.v_b = nT; nT += h.maxm;
  v.v_xl = nT; nT += h.maxm;
  v.ic_rs = nD; nD += h.icN;
  v.ic_cs = nD; nD += h.icN;
  v.v_den = nD; nD += n;
  v.v_col = nD; nD += std::max(1, std::max(h.maxm, h.icN));
  v.ic_piv = nI; nI += h.icN;
  v.nT = nT;
  v.nD = nD;
  v.nI = nI;
  return h;
}

}  // namespace kh
