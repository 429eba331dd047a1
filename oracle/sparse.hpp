// ORACLE — test infrastructure only. Not part of the product path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker.
//
// CPU restatement of kronspice's Gilbert–Peierls sparse LU
// (solver/sparse.go). Every function cites the Go lines it follows. Compiled
// with -ffp-contract=off so `x -= l*y` rounds twice, as the Go amd64 build does.
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <unordered_map>
#include <vector>
#include <algorithm>
#include <limits>

namespace oracle {

enum Err { OK = 0, SINGULAR = 1, NONFINITE = 2, BADSIZE = 3 };

// cmplx.Abs is math.Hypot (Go 1.24 math/hypot.go): the larger magnitude times
// sqrt(1 + (smaller/larger)^2), not a correctly rounded hypot.
inline double go_hypot(double p, double q) {
  p = std::fabs(p);
  q = std::fabs(q);
  if (std::isinf(p) || std::isinf(q)) return std::numeric_limits<double>::infinity();
  if (std::isnan(p) || std::isnan(q)) return std::numeric_limits<double>::quiet_NaN();
  if (p < q) std::swap(p, q);
  if (p == 0) return 0;
  q = q / p;
  return p * std::sqrt(1 + q * q);
}
inline double absval(double v) { return std::fabs(v); }
inline double absval(std::complex<double> v) { return go_hypot(v.real(), v.imag()); }  // cmplx.Abs

// Division as Go's complex128 division does it: Smith's algorithm (R. L.
// Smith, Algorithm 116, Commun. ACM 5(8):435, 1962), scaling by the larger
// part of the divisor. (Go's fix-up of the both-NaN results of infinite or
// zero operands is not needed here: every divisor is a pivot that passed the
// singularity test, and a non-finite quotient is caught downstream.)
inline double sdiv(double a, double b) { return a / b; }
inline std::complex<double> sdiv(std::complex<double> n, std::complex<double> m) {
  const double a = n.real(), b = n.imag(), c = m.real(), d = m.imag();
  double e, f;
  if (std::fabs(c) >= std::fabs(d)) {
    const double ratio = d / c;
    const double denom = c + ratio * d;
    e = (a + b * ratio) / denom;
    f = (b - a * ratio) / denom;
  } else {
    const double ratio = c / d;
    const double denom = d + ratio * c;
    e = (a * ratio + b) / denom;
    f = (b * ratio - a) / denom;
  }
  return {e, f};
}
inline bool finite(double v) { return !std::isnan(v) && !std::isinf(v); }
inline bool finite(std::complex<double> v) { return finite(v.real()) && finite(v.imag()); }

constexpr double kDiagPivotTol = 1e-3;  // solver/sparse.go:36
constexpr int kMaxSparse = 1 << 22;     // solver/sparse.go:39

// minimumDegree, solver/sparse.go:412-451: exact minimum degree on the
// off-diagonal pattern of A + A^T; ties go to the lowest index; the chosen
// vertex's neighbours are joined into a clique. Sets replace Go's maps (the
// result depends only on set contents, not on iteration order).
inline std::vector<int> minimumDegree(int n, const std::vector<int32_t>& ri, const std::vector<int32_t>& ci) {
  std::vector<std::vector<int32_t>> adj(n);  // kept sorted, unique
  auto ins = [&](int a, int b) {
    auto& v = adj[a];
    auto it = std::lower_bound(v.begin(), v.end(), b);
    if (it == v.end() || *it != b) v.insert(it, b);
  };
  auto del = [&](int a, int b) {
    auto& v = adj[a];
    auto it = std::lower_bound(v.begin(), v.end(), b);
    if (it != v.end() && *it == b) v.erase(it);
  };
  for (size_t k = 0; k < ri.size(); k++) {
    int i = ri[k], j = ci[k];
    if (i != j) { ins(i, j); ins(j, i); }
  }
  std::vector<char> gone(n, 0);
  std::vector<int> order;
  order.reserve(n);
  std::vector<int32_t> nbr;
  while ((int)order.size() < n) {
    int v = -1;
    size_t best = std::numeric_limits<size_t>::max();
    for (int i = 0; i < n; i++) {
      if (!gone[i] && adj[i].size() < best) { v = i; best = adj[i].size(); }
    }
    gone[v] = 1;
    order.push_back(v);
    nbr = adj[v];
    for (int32_t u : nbr) {
      del(u, v);
      for (int32_t w : nbr)
        if (w != u) ins(u, w);
    }
    adj[v].clear();
  }
  return order;
}

// sparseLU[T], solver/sparse.go:45-104.
template <typename T>
struct SparseLU {
  int n = 0;
  std::unordered_map<int64_t, int32_t> slot;
  std::vector<int32_t> ri, ci;
  std::vector<T> val;
  bool dirty = true;
  std::vector<int64_t> seqKey;
  std::vector<int32_t> seqSlot;
  int cursor = 0;
  std::vector<int32_t> q, qinv, ap, arow, aslot;
  std::vector<int32_t> lp, li, up, ui;
  std::vector<T> lx, ux, udiag;
  std::vector<int32_t> pinv, prow;
  std::vector<T> x;
  std::vector<int32_t> xi, pstack, mark;
  int32_t gen = 0;
  std::vector<T> w;
  bool valid = false;
  // Detail of the last singular error (pivot step and magnitude).
  int errStep = -1;
  double errBest = 0;
  int analyses = 0;

  // newSparseLU, solver/sparse.go:88-104
  int init(int n_) {
    if (n_ <= 0 || n_ > kMaxSparse) return BADSIZE;
    n = n_;
    q.assign(n, 0); qinv.assign(n, 0); ap.assign(n + 1, 0);
    lp.assign(n + 1, 0); up.assign(n + 1, 0); udiag.assign(n, T(0));
    pinv.assign(n, 0); prow.assign(n, 0);
    x.assign(n, T(0)); xi.assign(2 * n, 0); pstack.assign(n, 0); mark.assign(n, 0); w.assign(n, T(0));
    dirty = true;
    return OK;
  }

  // clear, solver/sparse.go:106-112
  void clear() {
    for (auto& v : val) v = T(0);
    valid = false;
    cursor = 0;
  }

  // add, solver/sparse.go:114-143
  void add(int i, int j, T v) {
    if (i < 0 || j < 0 || i >= n || j >= n) return;
    int64_t key = (int64_t)i * (int64_t)n + (int64_t)j;
    int c = cursor++;
    if (c < (int)seqKey.size() && seqKey[c] == key) {
      val[seqSlot[c]] += v;
      valid = false;
      return;
    }
    int32_t k;
    auto it = slot.find(key);
    if (it == slot.end()) {
      k = (int32_t)val.size();
      slot[key] = k;
      ri.push_back(i);
      ci.push_back(j);
      val.push_back(T(0));
      dirty = true;
    } else {
      k = it->second;
    }
    if (c < (int)seqKey.size()) {
      seqKey[c] = key;
      seqSlot[c] = k;
    } else {
      seqKey.push_back(key);
      seqSlot.push_back(k);
    }
    val[k] += v;
    valid = false;
  }

  T at(int i, int j) const {
    auto it = slot.find((int64_t)i * n + j);
    return it == slot.end() ? T(0) : val[it->second];
  }

  // analyze, solver/sparse.go:154-187
  void analyze() {
    analyses++;
    std::vector<int> order = minimumDegree(n, ri, ci);
    for (int k = 0; k < n; k++) {
      q[k] = order[k];
      qinv[order[k]] = k;
    }
    std::vector<int32_t> cnt(n + 1, 0);
    for (size_t k = 0; k < val.size(); k++) cnt[qinv[ci[k]] + 1]++;
    for (int k = 0; k < n; k++) cnt[k + 1] += cnt[k];
    ap = cnt;
    size_t nz = val.size();
    arow.assign(nz, 0);
    aslot.assign(nz, 0);
    std::vector<int32_t> next(cnt.begin(), cnt.begin() + n);
    for (size_t k = 0; k < val.size(); k++) {
      int c = qinv[ci[k]];
      int p = next[c]++;
      arow[p] = qinv[ri[k]];
      aslot[p] = (int32_t)k;
    }
    // Rows sorted ascending within each column (colSorter, :189-196). Rows are
    // distinct within a column, so the sort is unambiguous.
    for (int c = 0; c < n; c++) {
      int lo = ap[c], hi = ap[c + 1];
      std::vector<std::pair<int32_t, int32_t>> tmp;
      for (int p = lo; p < hi; p++) tmp.push_back({arow[p], aslot[p]});
      std::sort(tmp.begin(), tmp.end());
      for (int p = lo; p < hi; p++) { arow[p] = tmp[p - lo].first; aslot[p] = tmp[p - lo].second; }
    }
    dirty = false;
  }

  bool marked(int32_t r) const { return mark[r] == gen; }

  // reach, solver/sparse.go:301-352 (Davis cs_reach, iterative DFS)
  int reach(int k) {
    gen++;
    if (gen == std::numeric_limits<int32_t>::max()) {
      std::fill(mark.begin(), mark.end(), 0);
      gen = 1;
    }
    int top = n;
    int32_t* stack = xi.data();
    for (int p = ap[k]; p < ap[k + 1]; p++) {
      int32_t r = arow[p];
      if (mark[r] == gen) continue;
      int head = 0;
      stack[0] = r;
      while (head >= 0) {
        int32_t v = stack[head];
        int32_t j = pinv[v];
        if (mark[v] != gen) {
          mark[v] = gen;
          if (j >= 0) pstack[head] = lp[j];
        }
        bool done = true;
        if (j >= 0) {
          int32_t end = lp[j + 1];
          for (int32_t qq = pstack[head]; qq < end; qq++) {
            int32_t wv = li[qq];
            if (mark[wv] == gen) continue;
            pstack[head] = qq + 1;
            head++;
            stack[head] = wv;
            done = false;
            break;
          }
        }
        if (done) {
          head--;
          top--;
          xi[top] = v;
        }
      }
    }
    return top;
  }

  // factor, solver/sparse.go:199-293
  int factor() {
    if (dirty) analyze();
    double scale = 0.0;
    for (auto& v : val) {
      if (!finite(v)) return NONFINITE;
      double a = absval(v);
      if (a > scale) scale = a;
    }
    if (scale == 0) { errStep = -1; errBest = 0; return SINGULAR; }
    std::fill(pinv.begin(), pinv.end(), -1);
    li.clear(); lx.clear(); ui.clear(); ux.clear();
    for (int k = 0; k < n; k++) {
      lp[k] = (int32_t)li.size();
      up[k] = (int32_t)ui.size();
      int top = reach(k);
      for (int p = top; p < n; p++) x[xi[p]] = T(0);
      double colMax = 0.0;
      for (int p = ap[k]; p < ap[k + 1]; p++) {
        T v = val[aslot[p]];
        x[arow[p]] = v;
        double a = absval(v);
        if (a > colMax) colMax = a;
      }
      double tiny = colMax * (double)n * 1e-14;
      for (int p = top; p < n; p++) {
        int32_t r = xi[p];
        int32_t j = pinv[r];
        if (j < 0) continue;
        T xj = x[r];
        if (xj == T(0)) continue;
        for (int32_t qq = lp[j]; qq < lp[j + 1]; qq++) x[li[qq]] -= lx[qq] * xj;
      }
      int32_t piv = -1;
      double best = -1.0;
      for (int p = top; p < n; p++) {
        int32_t r = xi[p];
        if (pinv[r] >= 0) {
          ui.push_back(pinv[r]);
          ux.push_back(x[r]);
          continue;
        }
        double a = absval(x[r]);
        if (a > best || (a == best && r < piv)) { best = a; piv = r; }
      }
      if (piv < 0 || best <= tiny) {
        lp[k + 1] = (int32_t)li.size();
        up[k + 1] = (int32_t)ui.size();
        errStep = k;
        errBest = std::max(best, 0.0);
        return SINGULAR;
      }
      if (pinv[k] < 0 && marked(k) && absval(x[k]) >= kDiagPivotTol * best && absval(x[k]) > tiny) piv = k;
      T pv = x[piv];
      udiag[k] = pv;
      pinv[piv] = k;
      prow[k] = piv;
      for (int p = top; p < n; p++) {
        int32_t r = xi[p];
        if (pinv[r] < 0) {
          li.push_back(r);
          lx.push_back(sdiv(x[r], pv));
        }
      }
    }
    lp[n] = (int32_t)li.size();
    up[n] = (int32_t)ui.size();
    valid = true;
    return OK;
  }

  // solve, solver/sparse.go:355-401
  int solve(const T* b, T* xout) {
    if (!valid) {
      int e = factor();
      if (e != OK) return e;
    }
    for (int i = 0; i < n; i++) w[qinv[i]] = b[i];
    T* y = x.data();
    for (int k = 0; k < n; k++) {
      T yk = w[prow[k]];
      y[k] = yk;
      if (yk == T(0)) continue;
      for (int p = lp[k]; p < lp[k + 1]; p++) w[li[p]] -= lx[p] * yk;
    }
    for (int k = n - 1; k >= 0; k--) {
      T zk = sdiv(y[k], udiag[k]);
      y[k] = zk;
      if (zk == T(0)) continue;
      for (int p = up[k]; p < up[k + 1]; p++) y[ui[p]] -= ux[p] * zk;
    }
    for (int k = 0; k < n; k++) {
      T v = y[k];
      if (!finite(v)) return NONFINITE;
      xout[q[k]] = v;
    }
    return OK;
  }

  // AbsProduct, solver/sparse.go:481-490 (real) and :519-527 (complex)
  void absProduct(const T* xv, double* y) const {
    for (int i = 0; i < n; i++) y[i] = 0;
    for (size_t k = 0; k < val.size(); k++) y[ri[k]] += absval(val[k]) * absval(xv[ci[k]]);
  }

  // nnz, solver/sparse.go:405
  void nnz(int64_t* a, int64_t* l, int64_t* u) const {
    *a = (int64_t)val.size();
    *l = (int64_t)li.size();
    *u = (int64_t)ui.size() + n;
  }
};

// Dense and CDense, solver/dense.go:21-178 and solver/cdense.go:92-193
// (Doolittle LU with partial pivoting; complex magnitudes are cmplx.Abs).
template <typename T>
struct DenseT {
  int n = 0;
  std::vector<T> a, lu, y;
  std::vector<double> col;
  std::vector<int> piv;
  int errStep = -1;
  int init(int n_) {
    if (n_ <= 0 || n_ > 4096) return BADSIZE;  // MaxDense, :12
    n = n_;
    a.assign((size_t)n * n, 0);
    lu.assign((size_t)n * n, 0);
    piv.assign(n, 0);
    y.assign(n, 0);
    col.assign(n, 0);
    return OK;
  }
  void clear() { std::fill(a.begin(), a.end(), 0.0); }
  void add(int i, int j, T v) {
    if (i < 0 || j < 0 || i >= n || j >= n) return;
    a[(size_t)i * n + j] += v;
  }
  void absProduct(const T* x, double* yv) const {
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int j = 0; j < n; j++) s += absval(a[(size_t)i * n + j]) * absval(x[j]);
      yv[i] = s;
    }
  }
  int solve(const T* b, T* x) {
    lu = a;
    double scale = 0;
    for (const T& v : lu) {
      if (!finite(v)) return NONFINITE;
      double av = absval(v);
      if (av > scale) scale = av;
    }
    if (scale == 0) return SINGULAR;
    std::fill(col.begin(), col.end(), 0.0);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double av = absval(lu[(size_t)i * n + j]);
        if (av > col[j]) col[j] = av;
      }
    for (int k = 0; k < n; k++) {
      int p = k;
      double maxv = absval(lu[(size_t)k * n + k]);
      for (int i = k + 1; i < n; i++) {
        double v = absval(lu[(size_t)i * n + k]);
        if (v > maxv) { maxv = v; p = i; }
      }
      if (maxv <= col[k] * (double)n * 1e-14) { errStep = k; return SINGULAR; }
      piv[k] = p;
      if (p != k)
        for (int j = 0; j < n; j++) std::swap(lu[(size_t)k * n + j], lu[(size_t)p * n + j]);
      T pivot = lu[(size_t)k * n + k];
      for (int i = k + 1; i < n; i++) {
        T f = sdiv(lu[(size_t)i * n + k], pivot);
        lu[(size_t)i * n + k] = f;
        if (f == T(0)) continue;
        for (int j = k + 1; j < n; j++) lu[(size_t)i * n + j] -= f * lu[(size_t)k * n + j];
      }
    }
    for (int i = 0; i < n; i++) y[i] = b[i];
    for (int k = 0; k < n; k++)
      if (piv[k] != k) std::swap(y[k], y[piv[k]]);
    for (int i = 0; i < n; i++) {
      T s = y[i];
      for (int j = 0; j < i; j++) s -= lu[(size_t)i * n + j] * y[j];
      y[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
      T s = y[i];
      for (int j = i + 1; j < n; j++) s -= lu[(size_t)i * n + j] * x[j];
      x[i] = sdiv(s, lu[(size_t)i * n + i]);
    }
    for (int i = 0; i < n; i++)
      if (!finite(x[i])) return NONFINITE;
    return OK;
  }
};
using Dense = DenseT<double>;
using CDense = DenseT<std::complex<double>>;  // solver/cdense.go:92-193 (cmplx.Abs pivots)

}  // namespace oracle
