// ORACLE — test infrastructure only. Not part of the product path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker.
//
// CPU restatement of kron.JacobianTearing (kron/jacobian.go:66-712): a linear
// system solved by tearing. The unknowns fall into pieces (connected
// components of the pattern with the cut node pairs removed); with B the
// block-diagonal part and the coupling entries E = P·M·Q,
//   z = B⁻¹·b,  (I + M·Q·B⁻¹·P)·w = M·Q·z,  x = z − B⁻¹·P·w,
// every piece factored by an equilibrated dense LU and kept while its block is
// unchanged (or changed by at most ReuseTol of its largest entry), the
// solution refined against the whole matrix to a componentwise backward error
// of 1e-14 (accepted at 1e-12), else solved by an untorn LU (solver.Auto).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <limits>
#include <memory>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "sparse.hpp"

namespace oracle {

constexpr int kTearMaxRefine = 12;          // maxRefine, kron/diakoptics.go:83
constexpr double kTearRefineDone = 1e-14;   // refineDone, :84
constexpr double kTearRefineAccept = 1e-12; // refineAccept, :85

// math.Max: NaN if either argument is NaN, +0 over -0.
inline double go_max(double a, double b) {
  if (std::isinf(a) && a > 0) return a;
  if (std::isinf(b) && b > 0) return b;
  if (std::isnan(a) || std::isnan(b)) return std::numeric_limits<double>::quiet_NaN();
  if (a == 0 && a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}

// JacobianTearing's shared part (kron/jacobian.go:66-78): the cut node pairs
// and the counters every matrix it makes adds to.
struct Tearing {
  std::vector<std::pair<int, int>> cuts;  // node unknowns of each torn pair (-1: ground)
  double reuseTol = 0;
  int64_t solves = 0, factors = 0, reuses = 0, stale = 0, fallbacks = 0, refines = 0, maxPieces = 0;
};

// denseLU (kron/jacobian.go:591-712): equilibrated dense LU with partial
// pivoting (rows, then columns scaled to unit maximum).
template <typename T>
struct TearDense {
  int n = 0;
  std::vector<T> lu;
  std::vector<int> piv;
  std::vector<double> rs, cs, col;
  void init(int n_) {
    n = n_;
    lu.assign((size_t)n * n, T(0));
    piv.assign(n, 0);
    rs.assign(n, 0);
    cs.assign(n, 0);
    col.assign(n, 0);
  }
  struct Ent {
    int32_t i, j;
    T v;
  };
  // factor, :611-672
  int factor(const std::vector<Ent>& ent) {
    std::fill(lu.begin(), lu.end(), T(0));
    for (int i = 0; i < n; i++) rs[i] = cs[i] = col[i] = 0;
    for (const Ent& e : ent) rs[e.i] = go_max(rs[e.i], absval(e.v));
    for (int i = 0; i < n; i++) {
      if (rs[i] == 0) return SINGULAR;
      rs[i] = 1 / rs[i];
    }
    for (const Ent& e : ent) cs[e.j] = go_max(cs[e.j], absval(e.v) * rs[e.i]);
    for (int j = 0; j < n; j++) {
      if (cs[j] == 0) return SINGULAR;
      cs[j] = 1 / cs[j];
    }
    for (const Ent& e : ent) {
      const T s = e.v * T(rs[e.i]) * T(cs[e.j]);
      lu[(size_t)e.i * n + e.j] = lu[(size_t)e.i * n + e.j] + s;
    }
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) col[j] = go_max(col[j], absval(lu[(size_t)i * n + j]));
    for (int k = 0; k < n; k++) {
      int p = k;
      double maxv = absval(lu[(size_t)k * n + k]);
      for (int i = k + 1; i < n; i++) {
        const double v = absval(lu[(size_t)i * n + k]);
        if (v > maxv) { maxv = v; p = i; }
      }
      if (maxv <= col[k] * (double)n * 1e-14) return SINGULAR;
      piv[k] = p;
      if (p != k)
        for (int j = 0; j < n; j++) std::swap(lu[(size_t)k * n + j], lu[(size_t)p * n + j]);
      const T pivot = lu[(size_t)k * n + k];
      for (int i = k + 1; i < n; i++) {
        const T f = sdiv(lu[(size_t)i * n + k], pivot);
        lu[(size_t)i * n + k] = f;
        if (f == T(0)) continue;
        for (int j = k + 1; j < n; j++) {
          const T prod = f * lu[(size_t)k * n + j];
          lu[(size_t)i * n + j] = lu[(size_t)i * n + j] - prod;
        }
      }
    }
    return OK;
  }
  // solve, :676-712 (b and x differ)
  void solve(const T* b, T* x) const {
    for (int i = 0; i < n; i++) x[i] = b[i] * T(rs[i]);
    for (int k = 0; k < n; k++)
      if (piv[k] != k) std::swap(x[k], x[piv[k]]);
    for (int i = 0; i < n; i++) {
      T s = x[i];
      for (int j = 0; j < i; j++) {
        const T prod = lu[(size_t)i * n + j] * x[j];
        s = s - prod;
      }
      x[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
      T s = x[i];
      for (int j = i + 1; j < n; j++) {
        const T prod = lu[(size_t)i * n + j] * x[j];
        s = s - prod;
      }
      x[i] = sdiv(s, lu[(size_t)i * n + i]);
    }
    for (int j = 0; j < n; j++) x[j] = x[j] * T(cs[j]);
  }
};

// A record of one torn solve (for tests that replay an analysis's sequence
// of systems elsewhere): the matrix it was made on, its entries in slot
// order, values, right-hand side, solution and status. Values are (re, im)
// pairs when complex.
struct TearRecord {
  int mat = 0, cplx = 0, n = 0, status = 0;
  std::vector<int32_t> ri, ci;
  std::vector<double> val, b, x;
};
inline std::vector<TearRecord>* g_tear_rec = nullptr;  // recording when non-null
inline int g_tear_mats = 0;                            // matrices made so far (their ids)

inline void rec_push(std::vector<double>& out, double v) { out.push_back(v); }
inline void rec_push(std::vector<double>& out, std::complex<double> v) {
  out.push_back(v.real());
  out.push_back(v.imag());
}

// tornMat (kron/jacobian.go:199-589): one torn system; its pattern is learned
// from the stamps and the plan rebuilt when a new entry appears.
template <typename T>
struct TornMat {
  int n = 0;
  int id = 0;
  Tearing* jt = nullptr;
  std::unordered_map<int64_t, int32_t> slot;
  std::vector<int32_t> ri, ci;
  std::vector<T> val;
  int backend = 0;  // the untorn fallback: solver.NewMatrix(n, solver.Auto)

  struct Piece {
    std::vector<int32_t> idx, slots, li, lj, ports;
    TearDense<T> lu;
    std::vector<T> last;
    bool have = false;
    std::vector<std::vector<T>> W;
    std::vector<typename TearDense<T>::Ent> ent;
    std::vector<T> b, xl;
  };
  struct Plan {
    std::vector<std::unique_ptr<Piece>> pieces;
    std::vector<int32_t> pieceOf, local, off, rows, rowIdx;
    int icN = 0;
    TearDense<T> ic;
    std::vector<typename TearDense<T>::Ent> icEnt;
    std::vector<T> z, x, r, h, w, dx;
    std::vector<double> den;
  };
  std::unique_ptr<Plan> plan;
  bool fbSparse = false;
  SparseLU<T> fbS;
  DenseT<T> fbD;
  bool fbMade = false;

  void init(int n_, Tearing* t) {
    n = n_;
    jt = t;
    id = g_tear_mats++;
  }
  void clear() { std::fill(val.begin(), val.end(), T(0)); }  // :231-236
  void add(int i, int j, T v) {                                // :238-253
    if (i < 0 || j < 0 || i >= n || j >= n) return;
    const int64_t k = (int64_t)i * n + j;
    auto it = slot.find(k);
    int32_t s;
    if (it == slot.end()) {
      s = (int32_t)val.size();
      slot[k] = s;
      ri.push_back(i);
      ci.push_back(j);
      val.push_back(T(0));
      plan.reset();
    } else {
      s = it->second;
    }
    val[s] = val[s] + v;
  }
  void absProduct(const T* x, double* y) const {  // :255-262
    for (int i = 0; i < n; i++) y[i] = 0;
    for (size_t s = 0; s < val.size(); s++) y[ri[s]] += absval(val[s]) * absval(x[ci[s]]);
  }

  // buildPlan, :298-389
  int buildPlan() {
    std::unordered_map<int64_t, bool> cut;
    for (const auto& c : jt->cuts)
      if (c.first < n && c.second < n) {
        cut[(int64_t)c.first * n + c.second] = true;
        cut[(int64_t)c.second * n + c.first] = true;
      }
    std::vector<int32_t> parent(n);
    for (int i = 0; i < n; i++) parent[i] = i;
    auto find = [&](int32_t i) {
      while (parent[i] != i) {
        parent[i] = parent[parent[i]];
        i = parent[i];
      }
      return i;
    };
    for (size_t s = 0; s < val.size(); s++) {
      const int32_t i = ri[s], j = ci[s];
      if (i == j || cut.count((int64_t)i * n + j)) continue;
      const int32_t a = find(i), b = find(j);
      if (a != b) {
        if (a < b) parent[b] = a;
        else parent[a] = b;
      }
    }
    auto p = std::make_unique<Plan>();
    p->pieceOf.assign(n, 0);
    p->local.assign(n, 0);
    p->rowIdx.assign(n, -1);
    std::vector<int32_t> root(n, -1);
    for (int32_t i = 0; i < n; i++) {
      const int32_t r = find(i);
      if (root[r] < 0) {
        root[r] = (int32_t)p->pieces.size();
        p->pieces.push_back(std::make_unique<Piece>());
      }
      Piece& pc = *p->pieces[root[r]];
      p->pieceOf[i] = root[r];
      p->local[i] = (int32_t)pc.idx.size();
      pc.idx.push_back(i);
    }
    for (size_t s = 0; s < val.size(); s++) {
      const int32_t i = ri[s], j = ci[s];
      if (p->pieceOf[i] == p->pieceOf[j]) {
        Piece& pc = *p->pieces[p->pieceOf[i]];
        pc.slots.push_back((int32_t)s);
        pc.li.push_back(p->local[i]);
        pc.lj.push_back(p->local[j]);
        continue;
      }
      p->off.push_back((int32_t)s);
      if (p->rowIdx[i] < 0) {
        p->rowIdx[i] = (int32_t)p->rows.size();
        p->rows.push_back(i);
        p->pieces[p->pieceOf[i]]->ports.push_back(p->rowIdx[i]);
      }
    }
    for (auto& pcp : p->pieces) {
      Piece& pc = *pcp;
      const int nl = (int)pc.idx.size();
      if (nl > 4096) return BADSIZE;  // solver.MaxDense
      pc.lu.init(nl);
      pc.last.assign(pc.slots.size(), T(0));
      pc.ent.assign(pc.slots.size(), {0, 0, T(0)});
      pc.W.assign(pc.ports.size(), std::vector<T>(nl, T(0)));
      pc.b.assign(nl, T(0));
      pc.xl.assign(nl, T(0));
    }
    p->icN = (int)p->rows.size();
    if (p->icN > 0) p->ic.init(p->icN);
    p->z.assign(n, T(0));
    p->x.assign(n, T(0));
    p->r.assign(n, T(0));
    p->dx.assign(n, T(0));
    p->h.assign(p->icN, T(0));
    p->w.assign(p->icN, T(0));
    p->den.assign(n, 0);
    plan = std::move(p);
    if ((int64_t)plan->pieces.size() > jt->maxPieces) jt->maxPieces = (int64_t)plan->pieces.size();
    return OK;
  }

  int solve(const T* b, T* x) {
    const int e = solve_torn(b, x);
    if (g_tear_rec) {
      TearRecord r;
      r.mat = id;
      r.cplx = std::is_same<T, double>::value ? 0 : 1;
      r.n = n;
      r.status = e;
      r.ri = ri;
      r.ci = ci;
      for (const T& v : val) rec_push(r.val, v);
      for (int i = 0; i < n; i++) rec_push(r.b, b[i]);
      for (int i = 0; i < n; i++) rec_push(r.x, x[i]);
      g_tear_rec->push_back(std::move(r));
    }
    return e;
  }

  // Solve, :393-416
  int solve_torn(const T* b, T* x) {
    jt->solves++;
    for (const T& v : val) {
      const double f = absval(v);
      if (std::isnan(f) || std::isinf(f)) return NONFINITE;
    }
    if (!plan && buildPlan() != OK) return untorn(b, x);
    bool stale = false;
    bool ok = torn(b, x, true, &stale);
    if (!ok && stale) {
      bool dummy = false;
      ok = torn(b, x, false, &dummy);
    }
    if (!ok) return untorn(b, x);
    return OK;
  }

  // untorn, :418-432
  int untorn(const T* b, T* x) {
    jt->fallbacks++;
    if (!fbMade) {
      fbSparse = n >= 30;  // Backend.Resolve(Auto), solver/matrix.go:77-85
      if (fbSparse) fbS.init(n);
      else fbD.init(n);
      fbMade = true;
    }
    if (fbSparse) {
      fbS.clear();
      for (size_t s = 0; s < val.size(); s++) fbS.add(ri[s], ci[s], val[s]);
      return fbS.solve(b, x);
    }
    fbD.clear();
    for (size_t s = 0; s < val.size(); s++) fbD.add(ri[s], ci[s], val[s]);
    return fbD.solve(b, x);
  }

  // apply (the closure in torn, :503-544): out = J⁻¹·rhs through the pieces
  // and the interconnection.
  void apply(const T* rhs, T* out) {
    Plan& p = *plan;
    T* z = out;
    for (auto& pcp : p.pieces) {
      Piece& pc = *pcp;
      bool nz = false;
      for (size_t li = 0; li < pc.idx.size(); li++) {
        pc.b[li] = rhs[pc.idx[li]];
        nz = nz || rhs[pc.idx[li]] != T(0);
      }
      if (!nz) {
        for (int32_t g : pc.idx) z[g] = T(0);
        continue;
      }
      pc.lu.solve(pc.b.data(), pc.xl.data());
      for (size_t li = 0; li < pc.idx.size(); li++) z[pc.idx[li]] = pc.xl[li];
    }
    if (p.icN == 0) return;
    for (int a = 0; a < p.icN; a++) p.h[a] = T(0);
    for (int32_t s : p.off) {
      const T prod = val[s] * z[ci[s]];
      p.h[p.rowIdx[ri[s]]] = p.h[p.rowIdx[ri[s]]] + prod;
    }
    p.ic.solve(p.h.data(), p.w.data());
    for (auto& pcp : p.pieces) {
      Piece& pc = *pcp;
      for (size_t k = 0; k < pc.ports.size(); k++) {
        const T wk = p.w[pc.ports[k]];
        if (wk == T(0)) continue;
        const std::vector<T>& col = pc.W[k];
        for (size_t li = 0; li < pc.idx.size(); li++) {
          const T prod = wk * col[li];
          z[pc.idx[li]] = z[pc.idx[li]] - prod;
        }
      }
    }
  }

  // torn, :436-589
  bool torn(const T* b, T* x, bool allowStale, bool* usedStale) {
    Plan& p = *plan;
    *usedStale = false;
    for (auto& pcp : p.pieces) {
      Piece& pc = *pcp;
      bool same = pc.have;
      const bool small = pc.have && allowStale && jt->reuseTol > 0;
      double mx = 0, dmax = 0;
      for (size_t k = 0; k < pc.slots.size(); k++) {
        const T v = val[pc.slots[k]];
        if (v != pc.last[k]) {
          same = false;
          if (small) dmax = go_max(dmax, absval(v - pc.last[k]));
        }
        if (small) mx = go_max(mx, absval(v));
      }
      if (same) {
        jt->reuses++;
        continue;
      }
      if (small && dmax <= jt->reuseTol * mx) {
        jt->stale++;
        *usedStale = true;
        continue;
      }
      for (size_t k = 0; k < pc.slots.size(); k++) {
        pc.ent[k] = {pc.li[k], pc.lj[k], val[pc.slots[k]]};
        pc.last[k] = val[pc.slots[k]];
      }
      pc.have = false;
      if (pc.lu.factor(pc.ent) != OK) return false;
      jt->factors++;
      for (size_t k = 0; k < pc.ports.size(); k++) {
        std::fill(pc.b.begin(), pc.b.end(), T(0));
        pc.b[p.local[p.rows[pc.ports[k]]]] = T(1);
        pc.lu.solve(pc.b.data(), pc.W[k].data());
      }
      pc.have = true;
    }
    if (p.icN > 0) {
      p.icEnt.clear();
      for (int a = 0; a < p.icN; a++) p.icEnt.push_back({a, a, T(1)});
      for (int32_t s : p.off) {
        const int32_t r = ri[s], c = ci[s];
        const T v = val[s];
        if (v == T(0)) continue;
        Piece& pc = *p.pieces[p.pieceOf[c]];
        const int32_t lc = p.local[c];
        for (size_t k = 0; k < pc.ports.size(); k++) {
          const T g = pc.W[k][lc];
          if (g != T(0)) p.icEnt.push_back({p.rowIdx[r], pc.ports[k], v * g});
        }
      }
      if (p.ic.factor(p.icEnt) != OK) return false;
    }
    apply(b, x);
    double berr = std::numeric_limits<double>::infinity();
    for (int pass = 0; pass <= kTearMaxRefine; pass++) {
      for (int i = 0; i < n; i++) {
        p.r[i] = b[i];
        p.den[i] = absval(b[i]);
      }
      for (size_t s = 0; s < val.size(); s++) {
        const int32_t i = ri[s];
        const T t = val[s] * x[ci[s]];
        p.r[i] = p.r[i] - t;
        p.den[i] = p.den[i] + absval(t);
      }
      berr = 0;
      for (int i = 0; i < n; i++) {
        if (p.den[i] > 0) berr = go_max(berr, absval(p.r[i]) / p.den[i]);
        else if (p.r[i] != T(0)) berr = std::numeric_limits<double>::infinity();
      }
      if (berr <= kTearRefineDone || pass == kTearMaxRefine) break;
      apply(p.r.data(), p.dx.data());
      for (int i = 0; i < n; i++) x[i] = x[i] + p.dx[i];
      jt->refines++;
    }
    if (!(berr <= kTearRefineAccept)) return false;
    for (int i = 0; i < n; i++) {
      const double f = absval(x[i]);
      if (std::isnan(f) || std::isinf(f)) return false;
    }
    return true;
  }
};

}  // namespace oracle
