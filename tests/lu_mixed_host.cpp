// Host build of the mixed-precision LU (kronspice_amd/csrc/kh_lu.h, compiled
// for the CPU with KH_HD empty) against the double-precision LU of the same
// header. Prints one line per check; exit status 0 when every check holds.
//  1. real MNA systems (conductance grids with voltage sources, resistor
//     ladders, random unsymmetric systems): every certified factorization has
//     the FP64 pivot rows; certified solutions within 1e-10 of FP64's.
//  2. complex AC systems of an RLC grid, 128 lanes in structure-of-arrays
//     layout, phases interleaved as concurrent wavefronts run them (half the
//     lanes refactor in complex128 before the other half runs its mixed path):
//     every lane's solution within 1e-12 of a fresh complex128 solve.
#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <vector>

#include "kh_symbolic.h"

using namespace kh;

namespace {

template <typename T>
struct Work {
  std::vector<T> xw, w, ud, lx, ux, x;
  std::vector<int32_t> li, ui, lp, up, pinv, prow, xi, ps, mark, err, rch, rp, dk, cpid, pid;
  LuArgsT<T> a = {};
  Work(int n, int64_t cap, int B, const PatternPool& pool, const T* val, const T* rhs, int nslots) {
    xw.assign((size_t)n * B, T{}); w = xw; ud = xw; x = xw;
    lx.assign((size_t)cap * B, T{}); ux = lx;
    li.assign((size_t)cap * B, 0); ui = li;
    lp.assign((size_t)(n + 1) * B, 0); up = lp; rp = lp;
    pinv.assign((size_t)n * B, 0); prow = pinv; xi = pinv; ps = pinv; mark = pinv; dk = pinv;
    err.assign(B, 0); cpid.assign(B, -1); pid.assign(B, 0);
    rch.assign((size_t)(2 * cap + n) * B, 0);
    a.n = n; a.nslots = nslots; a.B = B; a.capL = cap; a.capU = cap; a.capR = 2 * cap + n;
    a.pats = pool.desc.data(); a.pool = pool.pool.data(); a.pid = pid.data(); a.val = val; a.rhs = rhs;
    a.x = x.data(); a.xw = xw.data(); a.w = w.data(); a.udiag = ud.data(); a.lx = lx.data(); a.ux = ux.data();
    a.li = li.data(); a.ui = ui.data(); a.lp = lp.data(); a.up = up.data(); a.pinv = pinv.data();
    a.prow = prow.data(); a.xi = xi.data(); a.pstack = ps.data(); a.mark = mark.data(); a.err_step = err.data();
    a.rch = rch.data(); a.rp = rp.data(); a.dk = dk.data(); a.cpid = cpid.data();
  }
};

struct Entries {
  int n = 0;
  std::map<std::pair<int, int>, int> slot;
  std::vector<int32_t> ri, ci;
  int at(int i, int j) {
    auto f = slot.find({i, j});
    if (f != slot.end()) return f->second;
    slot[{i, j}] = (int)ri.size();
    ri.push_back(i);
    ci.push_back(j);
    return (int)ri.size() - 1;
  }
};

double lognu(std::mt19937_64& g, double a, double b) {
  return std::exp(std::log(a) + std::uniform_real_distribution<double>(0, 1)(g) * (std::log(b) - std::log(a)));
}

// Check 1: one real system; returns 0 uncertified, 1 certified and equal, -1 a violation.
int real_case(Entries& E, std::vector<double>& v, std::mt19937_64& g, double tol) {
  const int n = E.n;
  std::vector<int32_t> slots(E.ri.size());
  for (size_t k = 0; k < slots.size(); k++) slots[k] = (int32_t)k;
  PatternPool pool;
  pool.add(analyze_pattern(n, E.ri, E.ci, slots), "p");
  const int64_t cap = 2 * pool.max_fill + 4 * n + 64;
  std::vector<double> b(n);
  for (auto& x : b) x = std::uniform_real_distribution<double>(-1, 1)(g);
  Work<double> A(n, cap, 1, pool, v.data(), b.data(), (int)v.size()), M(n, cap, 1, pool, v.data(), b.data(), (int)v.size());
  int s = lu_factor<double, true>(A.a, 0);
  if (s == LU_OK) s = lu_solve<double, true>(A.a, 0);
  if (s != LU_OK) return 0;
  int it = 0;
  int m = lu_factor_mixed<double, true>(M.a, 0);
  if (m == LU_OK) m = lu_solve_mixed<double, true>(M.a, 0, &it);
  if (m != LU_OK) return 0;
  if (A.prow != M.prow) return -1;
  double xm = 0, d = 0;
  for (int i = 0; i < n; i++) {
    xm = std::max(xm, std::fabs(A.x[i]));
    d = std::max(d, std::fabs(A.x[i] - M.x[i]));
  }
  return d <= tol * xm ? 1 : -1;
}

}  // namespace

int main() {
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> U(0, 1);
  int bad = 0;
  // 1. real systems
  int cert[3] = {0, 0, 0}, tot[3] = {0, 0, 0};
  for (int t = 0; t < 600; t++) {
    const int kind = t % 3;
    Entries E;
    std::map<int, double> acc;
    std::vector<double> v;
    auto add = [&](int i, int j, double x) { acc[E.at(i, j)] += x; };
    if (kind == 0) {  // conductance grid, gmin, voltage sources
      const int w = 3 + t % 9, h = 3 + (t / 3) % 11, ns = 1 + t % 3, nn = w * h;
      E.n = nn + ns;
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
          const int a = y * w + x;
          for (int b : {x + 1 < w ? a + 1 : -1, y + 1 < h ? a + w : -1}) {
            if (b < 0) continue;
            const double gg = lognu(g, 1e-4, 1.0);
            add(a, a, gg); add(b, b, gg); add(a, b, -gg); add(b, a, -gg);
          }
          add(a, a, 1e-12);
        }
      for (int k = 0; k < ns; k++) {
        const int node = (k * 7919) % nn;
        add(node, nn + k, 1); add(nn + k, node, 1);
      }
    } else if (kind == 1) {  // resistor ladder driven by a source
      const int nodes = 64;
      E.n = nodes + 1;
      for (int k = 0; k < nodes; k++) {
        add(k, k, 1 / lognu(g, 100, 1e5));
        if (k + 1 < nodes) {
          const double gg = 1 / lognu(g, 100, 1e5);
          add(k, k, gg); add(k + 1, k + 1, gg); add(k, k + 1, -gg); add(k + 1, k, -gg);
        }
      }
      add(0, nodes, 1); add(nodes, 0, 1);
    } else {  // random unsymmetric
      const int n = 10 + t % 50;
      E.n = n;
      for (int i = 0; i < n; i++) add(i, i, (U(g) - 0.5) * 4);
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
          if (i != j && U(g) < 0.08) add(i, j, (U(g) - 0.5) * std::exp(U(g) * 6 - 3));
    }
    v.assign(E.ri.size(), 0.0);
    for (auto& kv : acc) v[kv.first] = kv.second;
    const int r = real_case(E, v, g, kind == 1 ? 1e-12 : 1e-10);
    tot[kind]++;
    if (r > 0) cert[kind]++;
    if (r < 0) bad++;
  }
  printf("real: certified grids %d/%d, ladders %d/%d, random %d/%d, violations %d\n", cert[0], tot[0], cert[1], tot[1],
         cert[2], tot[2], bad);
  if (cert[0] < tot[0] * 9 / 10 || cert[1] < tot[1] * 9 / 10 || cert[2] == 0) bad++;

  // 2. complex SoA lanes of an RLC grid, interleaved phases
  const int w = 8, h = 8, D = 4, NF = 32, BB = D * NF, nn = w * h, ne = (w - 1) * h + w * (h - 1);
  Entries E;
  E.n = nn + 2 * ne;
  struct St { int s, kind, e; double sign; };
  std::vector<St> st;
  int e = 0;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int a = y * w + x;
      for (int b : {x + 1 < w ? a + 1 : -1, y + 1 < h ? a + w : -1}) {
        if (b < 0) continue;
        const int mid = nn + e, br = nn + ne + e;
        st.push_back({E.at(a, a), 0, e, 1}); st.push_back({E.at(mid, mid), 0, e, 1});
        st.push_back({E.at(a, mid), 0, e, -1}); st.push_back({E.at(mid, a), 0, e, -1});
        st.push_back({E.at(mid, br), 1, e, 1}); st.push_back({E.at(b, br), 1, e, -1});
        st.push_back({E.at(br, mid), 1, e, 1}); st.push_back({E.at(br, b), 1, e, -1});
        st.push_back({E.at(br, br), 2, e, 1});
        e++;
      }
      st.push_back({E.at(a, a), 3, a, 1});
    }
  st.push_back({E.at(0, 0), 4, 0, 1});
  std::vector<std::vector<double>> G(D, std::vector<double>(ne)), Lv = G, Cv(D, std::vector<double>(nn));
  for (int d = 0; d < D; d++) {
    for (int k = 0; k < ne; k++) { G[d][k] = 1 / lognu(g, 0.1, 10); Lv[d][k] = lognu(g, 1e-9, 1e-6); }
    for (int k = 0; k < nn; k++) Cv[d][k] = lognu(g, 1e-12, 1e-9);
  }
  const int nnz = (int)E.ri.size(), n = E.n;
  std::vector<kcplx> vv((size_t)nnz * BB, kcplx{0, 0}), b((size_t)n * BB, kcplx{0, 0});
  for (int l = 0; l < BB; l++) {
    const int d = l % D, f = l / D;
    const double om = 2 * M_PI * std::pow(10.0, 3 + 7.0 * f / (NF - 1));
    for (const auto& q : st) {
      kcplx x = q.kind == 0 ? kcplx{q.sign * G[d][q.e], 0} : q.kind == 1 ? kcplx{q.sign, 0}
              : q.kind == 2 ? kcplx{0, -om * Lv[d][q.e]} : q.kind == 3 ? kcplx{0, om * Cv[d][q.e]} : kcplx{1e-6, 0};
      vv[(size_t)q.s * BB + l] = vv[(size_t)q.s * BB + l] + x;
    }
    b[l] = {1, 0};
  }
  std::vector<int32_t> slots(nnz);
  for (int k = 0; k < nnz; k++) slots[k] = k;
  PatternPool pool;
  pool.add(analyze_pattern(n, E.ri, E.ci, slots), "p");
  const int64_t cap = 2 * pool.max_fill + 4 * n + 64;
  Work<kcplx> A(n, cap, BB, pool, vv.data(), b.data(), nnz), M(n, cap, BB, pool, vv.data(), b.data(), nnz);
  for (int l = 0; l < BB; l++) {
    int s = lu_factor<kcplx>(A.a, l);
    if (s == LU_OK) s = lu_solve<kcplx>(A.a, l);
    if (s != LU_OK) bad++;
  }
  std::vector<int> redo(BB, 0);
  int ccert = 0;
  auto mixed = [&](int l) {
    int it = 0, s = lu_factor_mixed<kcplx>(M.a, l);
    if (s == LU_OK) s = lu_solve_mixed<kcplx>(M.a, l, &it);
    if (s == LU_OK) ccert++; else redo[l] = 1;
  };
  for (int l = 0; l < BB / 2; l++) {
    mixed(l);
    if (redo[l]) lu_factor<kcplx>(M.a, l);
  }
  for (int l = BB / 2; l < BB; l++) mixed(l);
  for (int l = 0; l < BB / 2; l++)
    if (redo[l]) lu_solve<kcplx>(M.a, l);
  for (int l = BB / 2; l < BB; l++)
    if (redo[l] && lu_factor<kcplx>(M.a, l) == LU_OK) lu_solve<kcplx>(M.a, l);
  int cbad = 0;
  for (int l = 0; l < BB; l++) {
    double xm = 0, d = 0;
    for (int i = 0; i < n; i++) {
      const kcplx p = A.x[(size_t)i * BB + l], q = M.x[(size_t)i * BB + l];
      xm = std::max(xm, std::hypot(p.re, p.im));
      d = std::max(d, std::hypot(p.re - q.re, p.im - q.im));
    }
    if (!(d <= 1e-12 * xm)) cbad++;
    if (redo[l] && d != 0) cbad++;  // a redone lane is the complex128 path's own result
  }
  printf("complex: %d lanes, certified %d, wrong %d\n", BB, ccert, cbad);
  if (cbad || ccert == 0) bad++;
  return bad ? 1 : 0;
}
