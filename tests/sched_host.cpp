// Host build of the compiled-schedule LU (kronspice_amd/csrc/kh_sched.h,
// compiled for the CPU with KH_HD empty) against the per-lane LU of kh_lu.h.
// Prints one line per case; exit status 0 when every check holds:
//  * sched_lu returns LU_OK exactly for the systems whose own factorization
//    (lu_factor) succeeds with the schedule's pivot rows and whose solution is
//    finite, and then its solution, U diagonal, L and U values are
//    bit-identical to lu_factor + lu_solve's;
//  * LU_UNSURE otherwise (another pivot row, a singular or non-finite system),
//    leaving the system to the per-lane path;
//  * the compact-working-column form (sched_lu_lds, three lanes sharing one
//    block, private arrays blocked by four lanes) agrees with sched_lu: status,
//    departure column and row, solution.
// Systems: conductance grids with voltage sources (zero diagonals: the
// diagonal preference decides), complex RLC grid admittances at frequencies
// across five decades, systems whose pivot rows differ from the probe's
// (a column's diagonal made small), exact zeros on the reach (the skipped
// updates), singular and non-finite systems, magnitude ties.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#include "kh_sched.h"
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

bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0 || (std::isnan(a) && std::isnan(b)); }
bool same(kcplx a, kcplx b) { return same(a.re, b.re) && same(a.im, b.im); }

// A system family: pattern (entries), a value generator per system.
template <typename T>
struct Family {
  Entries E;
  std::vector<std::vector<T>> val;  // [B][nslots]
  std::vector<std::vector<T>> rhs;  // [B][n]
};

// Conductance grid w x h with voltage sources on a few nodes (branch rows with
// zero diagonals), conductances log-uniform over `decades`.
Family<double> grid_family(int w, int h, int nv, int B, double decades, uint64_t seed) {
  Family<double> F;
  std::mt19937_64 g(seed);
  const int nn = w * h, n = nn + nv;
  F.E.n = n;
  struct Stamp { int s; int kind; int e; };
  std::vector<std::pair<int, int>> edges;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      if (x + 1 < w) edges.push_back({y * w + x, y * w + x + 1});
      if (y + 1 < h) edges.push_back({y * w + x, (y + 1) * w + x});
    }
  std::vector<int> vnode(nv);
  for (int k = 0; k < nv; k++) vnode[k] = (int)((k * 7919 + 13) % nn);
  // structure: every edge's 2x2, a shunt on every node, each source's branch
  for (auto& e : edges) { F.E.at(e.first, e.first); F.E.at(e.second, e.second); F.E.at(e.first, e.second); F.E.at(e.second, e.first); }
  for (int i = 0; i < nn; i++) F.E.at(i, i);
  for (int k = 0; k < nv; k++) { F.E.at(vnode[k], nn + k); F.E.at(nn + k, vnode[k]); }
  const int ns = (int)F.E.ri.size();
  for (int b = 0; b < B; b++) {
    std::vector<double> v(ns, 0.0), r(n, 0.0);
    for (auto& e : edges) {
      const double gg = lognu(g, 1.0, std::pow(10.0, decades));
      v[F.E.at(e.first, e.first)] += gg; v[F.E.at(e.second, e.second)] += gg;
      v[F.E.at(e.first, e.second)] -= gg; v[F.E.at(e.second, e.first)] -= gg;
    }
    for (int i = 0; i < nn; i++) v[F.E.at(i, i)] += lognu(g, 1e-3, 1e-1);
    for (int k = 0; k < nv; k++) { v[F.E.at(vnode[k], nn + k)] += 1; v[F.E.at(nn + k, vnode[k])] += 1; r[nn + k] = 1 + k; }
    for (int i = 0; i < nn; i++) r[i] = (i % 5 == 0) ? lognu(g, 1e-3, 1.0) : 0.0;
    F.val.push_back(v);
    F.rhs.push_back(r);
  }
  return F;
}

// Complex admittance of an RLC grid (series R + jwL edges, shunt jwC, one
// resistor to ground at node 0), one system per frequency.
Family<kcplx> rlc_family(int w, int h, int B, uint64_t seed) {
  Family<kcplx> F;
  std::mt19937_64 g(seed);
  const int nn = w * h;
  F.E.n = nn;
  std::vector<std::pair<int, int>> edges;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      if (x + 1 < w) edges.push_back({y * w + x, y * w + x + 1});
      if (y + 1 < h) edges.push_back({y * w + x, (y + 1) * w + x});
    }
  for (auto& e : edges) { F.E.at(e.first, e.first); F.E.at(e.second, e.second); F.E.at(e.first, e.second); F.E.at(e.second, e.first); }
  for (int i = 0; i < nn; i++) F.E.at(i, i);
  std::vector<double> R(edges.size()), L(edges.size()), Cc(nn);
  for (auto& r : R) r = lognu(g, 0.1, 10);
  for (auto& l : L) l = lognu(g, 1e-9, 1e-6);
  for (auto& c : Cc) c = lognu(g, 1e-12, 1e-9);
  const int ns = (int)F.E.ri.size();
  for (int b = 0; b < B; b++) {
    const double om = 2 * M_PI * std::pow(10.0, 3.0 + 5.0 * b / std::max(1, B - 1));
    std::vector<kcplx> v(ns, kcplx{0, 0}), r(nn, kcplx{0, 0});
    for (size_t k = 0; k < edges.size(); k++) {
      const kcplx z{R[k], om * L[k]};
      const kcplx y = lu_div(kcplx{1, 0}, z);
      auto& e = edges[k];
      v[F.E.at(e.first, e.first)] = v[F.E.at(e.first, e.first)] + y;
      v[F.E.at(e.second, e.second)] = v[F.E.at(e.second, e.second)] + y;
      v[F.E.at(e.first, e.second)] = v[F.E.at(e.first, e.second)] - y;
      v[F.E.at(e.second, e.first)] = v[F.E.at(e.second, e.first)] - y;
    }
    for (int i = 0; i < nn; i++) v[F.E.at(i, i)] = v[F.E.at(i, i)] + kcplx{0, om * Cc[i]};
    v[F.E.at(0, 0)] = v[F.E.at(0, 0)] + kcplx{1e-6, 0};
    r[0] = kcplx{1, 0};
    F.val.push_back(v);
    F.rhs.push_back(r);
  }
  return F;
}

// The MNA form of an RLC mesh: every edge a resistor to an internal node and
// an inductor branch (its current an unknown, its row V_m - V_b - jwL I = 0),
// a capacitor from every node to ground, 1 Mohm from node 0 to ground; one
// system per frequency, log-spaced over seven decades. The inductor branch
// columns pivot off the diagonal at low frequencies (|jwL| below the KLU
// threshold against the +-1 incidence entries) and on it at high ones: the
// pivot sequence changes in bands of frequency.
Family<kcplx> rlc_mna_family(int w, int h, int B, uint64_t seed) {
  Family<kcplx> F;
  std::mt19937_64 g(seed);
  const int nn = w * h;
  std::vector<std::pair<int, int>> edges;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      if (x + 1 < w) edges.push_back({y * w + x, y * w + x + 1});
      if (y + 1 < h) edges.push_back({y * w + x, (y + 1) * w + x});
    }
  const int ne = (int)edges.size(), n = nn + 2 * ne;  // nodes, internal nodes, branch currents
  F.E.n = n;
  auto mid = [&](int e) { return nn + e; };
  auto br = [&](int e) { return nn + ne + e; };
  for (int i = 0; i < nn; i++) F.E.at(i, i);
  for (int e = 0; e < ne; e++) {
    const int a = edges[e].first, b = edges[e].second, m = mid(e), k = br(e);
    F.E.at(a, a); F.E.at(m, m); F.E.at(a, m); F.E.at(m, a);
    F.E.at(m, k); F.E.at(b, k); F.E.at(k, m); F.E.at(k, b); F.E.at(k, k);
  }
  std::vector<double> R(ne), L(ne), Cc(nn);
  for (auto& r : R) r = lognu(g, 0.1, 10);
  for (auto& l : L) l = lognu(g, 1e-9, 1e-6);
  for (auto& c : Cc) c = lognu(g, 1e-12, 1e-9);
  const int ns = (int)F.E.ri.size();
  for (int b = 0; b < B; b++) {
    const double om = 2 * M_PI * std::pow(10.0, 3.0 + 7.0 * b / std::max(1, B - 1));
    std::vector<kcplx> v(ns, kcplx{0, 0}), r(n, kcplx{0, 0});
    auto add = [&](int i, int j, kcplx y) { v[F.E.at(i, j)] = v[F.E.at(i, j)] + y; };
    for (int i = 0; i < nn; i++) add(i, i, kcplx{0, om * Cc[i]});
    add(0, 0, kcplx{1e-6, 0});
    for (int e = 0; e < ne; e++) {
      const int a = edges[e].first, bb = edges[e].second, m = mid(e), k = br(e);
      const double gg = 1 / R[e];
      add(a, a, kcplx{gg, 0}); add(m, m, kcplx{gg, 0}); add(a, m, kcplx{-gg, 0}); add(m, a, kcplx{-gg, 0});
      add(m, k, kcplx{1, 0}); add(bb, k, kcplx{-1, 0}); add(k, m, kcplx{1, 0}); add(k, bb, kcplx{-1, 0});
      add(k, k, kcplx{0, -om * L[e]});
    }
    r[0] = kcplx{1, 0};
    F.val.push_back(v);
    F.rhs.push_back(r);
  }
  return F;
}

int failures = 0;

// Runs every system of F through sched_lu (schedule from system `probe`) and
// through lu_factor + lu_solve; checks the contract. Returns the number of
// systems sched_lu accepted.
template <typename T>
int check(const char* name, const Family<T>& F, int probe, int expect_ok_min, int expect_ok_max) {
  const int n = F.E.n, B = (int)F.val.size(), ns = (int)F.E.ri.size();
  PatternPool pool;
  std::vector<int32_t> slots(ns);
  for (int s = 0; s < ns; s++) slots[s] = s;
  Pattern pat = analyze_pattern(n, F.E.ri, F.E.ci, slots);
  pool.add(pat, "p");
  const int64_t cap = 2 * pool.max_fill + 4 * (int64_t)n + 64;
  // structure-of-arrays copies
  std::vector<T> val((size_t)ns * B), rhs((size_t)n * B);
  for (int b = 0; b < B; b++) {
    for (int s = 0; s < ns; s++) val[(size_t)s * B + b] = F.val[b][s];
    for (int i = 0; i < n; i++) rhs[(size_t)i * B + b] = F.rhs[b][i];
  }
  SchedHost sh;
  if (!build_schedule<T>(pool.desc, pool.pool, 0, n, ns, F.val[probe].data(), cap, sh)) {
    printf("%s: probe does not factor\n", name);
    failures++;
    return 0;
  }
  SchedView v;
  v.n = n; v.pid = 0;
  v.prow = sh.prow.data(); v.lp = sh.lp.data(); v.li = sh.li.data(); v.up = sh.up.data(); v.ue = sh.ue.data();
  v.dc = sh.dc.data();
  v.mmax = sh.mmax; v.tbase = sh.tbase.data(); v.tpos = sh.tpos.data(); v.apos = sh.apos.data();
  v.dslot = sh.dslot.data();
  const PatDesc& pd = pool.desc[0];
  v.ap = pool.pool.data() + pd.ap; v.arow = pool.pool.data() + pd.arow; v.aslot = pool.pool.data() + pd.aslot;
  v.q = pool.pool.data() + pd.q; v.qinv = pool.pool.data() + pd.qinv;
  Work<T> S(n, cap, B, pool, val.data(), rhs.data(), ns), P(n, cap, B, pool, val.data(), rhs.data(), ns),
      Q(n, cap, B, pool, val.data(), rhs.data(), ns), R(n, cap, B, pool, val.data(), rhs.data(), ns);
  // the compact form: 3 lanes sharing one working-column block (slot * 3 + lane)
  std::vector<double> X((size_t)sh.mmax * 3 * (sizeof(T) / sizeof(double)));
  int lds_diff = 0;
  int ok = 0, bad = 0, unsure = 0, redo_diff_piv = 0, redo_fail = 0;
  for (int c = 0; c < B; c++) {
    int d0 = 0, d1 = 0, p0 = 0, p1 = 0;
    const int ss = sched_lu<T>(v, S.a, c, sched_lane_soa(S.a, c), &d0, &p0);
    {
      // the device-memory form with the private arrays blocked by two lanes
      int d2 = 0, p2 = 0;
      const int sb = sched_lu<T>(v, R.a, c, sched_lane_blocked(R.a, c, 2), &d2, &p2);
      bool eq = sb == ss && (ss != LU_UNSURE || (d0 == d2 && (d0 < 0 || p0 == p2)));
      if (ss == LU_OK)
        for (int i = 0; i < n && eq; i++) eq = same(S.x[(size_t)i * B + c], R.x[(size_t)i * B + c]);
      if (!eq) lds_diff++;
    }
    {
      // the compact form: same status, departure column and row, solution
      const int sq = sched_lu_lds<T>(v, Q.a, c, sched_lane_blocked(Q.a, c, 4), X.data(), 3, c % 3, &d1, &p1);
      bool eq = sq == ss;
      if (sq == LU_UNSURE && ss == LU_UNSURE) eq = eq && d0 == d1 && (d0 < 0 || p0 == p1);
      if (ss == LU_OK)
        for (int i = 0; i < n && eq; i++) eq = same(S.x[(size_t)i * B + c], Q.x[(size_t)i * B + c]);
      if (!eq) lds_diff++;
    }
    int sp = lu_factor<T>(P.a, c);
    bool piv_same = sp == LU_OK;
    if (sp == LU_OK)
      for (int k = 0; k < n && piv_same; k++) piv_same = P.prow[(size_t)k * B + c] == sh.prow[k];
    if (sp == LU_OK) sp = lu_solve<T>(P.a, c);
    if (ss == LU_OK) {
      ok++;
      bool eq = piv_same && sp == LU_OK;
      for (int i = 0; i < n && eq; i++) eq = same(S.x[(size_t)i * B + c], P.x[(size_t)i * B + c]);
      for (int k = 0; k < n && eq; k++) eq = same(S.ud[(size_t)k * B + c], P.ud[(size_t)k * B + c]);
      for (int64_t t = 0; t < (int64_t)sh.li.size() && eq; t++) eq = same(S.lx[t * B + c], P.lx[t * B + c]);
      for (int64_t t = 0; t < (int64_t)sh.ue.size() && eq; t++) eq = same(S.ux[t * B + c], P.ux[t * B + c]);
      if (!eq) {
        bad++;
        if (bad <= 3) printf("%s: system %d accepted but differs from the per-lane LU (pivots same %d)\n", name, c, piv_same);
      }
    } else {
      unsure++;
      if (ss != LU_UNSURE) { bad++; printf("%s: system %d status %d\n", name, c, ss); }
      if (piv_same && sp == LU_OK) {
        bad++;
        if (bad <= 3) printf("%s: system %d refused though its pivots follow the schedule\n", name, c);
      }
      if (sp != LU_OK) redo_fail++;
      else redo_diff_piv++;
    }
  }
  const bool range = ok >= expect_ok_min && ok <= expect_ok_max;
  if (lds_diff) {
    printf("%s: the compact-column or blocked form differs from the structure-of-arrays form on %d systems\n", name, lds_diff);
    bad++;
  }
  printf("%s: n %d, systems %d, lockstep %d, refused %d (other pivots %d, failing %d), nnzL %zu nnzU %zu%s%s\n",
         name, n, B, ok, unsure, redo_diff_piv, redo_fail, sh.li.size(), sh.ue.size(), bad ? " MISMATCH" : "",
         range ? "" : " UNEXPECTED COUNT");
  if (bad || !range) failures++;
  return ok;
}

}  // namespace

int main() {
  // 1. real grids with voltage sources (their branch columns pivot off the
  //    diagonal, on whichever candidate is largest: systems differ)
  {
    auto F = grid_family(12, 9, 4, 48, 2.0, 1);
    check("grid12x9 (2 decades)", F, 0, 1, 48);
  }
  // 2. four decades of conductance: pivots may move; the contract still holds
  {
    auto F = grid_family(10, 10, 6, 64, 6.0, 2);
    check("grid10x10 (6 decades)", F, 0, 1, 64);
  }
  // 3. pivot rows that differ from the probe's: some systems get one node's
  //    shunt and edges scaled so far down that the diagonal preference fails
  {
    auto F = grid_family(8, 8, 3, 32, 1.0, 3);
    for (int b = 1; b < 32; b += 3) {
      // a voltage-source node's diagonal and its edges made tiny: its branch
      // row wins the pivot instead of the diagonal
      const int vn = (int)((0 * 7919 + 13) % 64);
      for (int s = 0; s < (int)F.E.ri.size(); s++)
        if (F.E.ri[s] == vn || F.E.ci[s] == vn)
          if (F.E.ri[s] < 64 && F.E.ci[s] < 64) F.val[b][s] *= 1e-9;
    }
    const int ok = check("grid8x8 (pivot changes)", F, 0, 1, 31);
    if (ok == 32) { printf("pivot-change case: no system refused\n"); failures++; }
  }
  // 4. exact zeros in the right-hand side and on the reach (skipped updates),
  //    a singular system (an all-zero column), a non-finite value
  {
    auto F = grid_family(6, 6, 2, 16, 1.0, 4);
    for (auto& r : F.rhs) for (size_t i = 0; i < r.size(); i++) if (i % 3) r[i] = 0.0;
    for (int s = 0; s < (int)F.E.ri.size(); s++) if (F.E.ci[s] == 5) F.val[3][s] = 0.0;
    F.val[7][2] = NAN;
    F.val[9][0] = INFINITY;
    check("grid6x6 (zeros, singular, non-finite)", F, 0, 1, 13);
  }
  // 5. ties: equal magnitudes in a column (unit conductances, no shunt jitter)
  {
    auto F = grid_family(5, 5, 2, 8, 0.0, 5);
    check("grid5x5 (ties)", F, 0, 8, 8);
  }
  // 6. complex RLC grid admittances over five decades of frequency, the
  //    probe at the lowest frequency and at the highest
  {
    auto F = rlc_family(10, 10, 64, 6);
    check("rlc10x10 (probe 1 kHz)", F, 0, 64, 64);
    check("rlc10x10 (probe 100 MHz)", F, 63, 64, 64);
  }
  // 7. the MNA form of an RLC mesh over seven decades: schedules from probes
  //    at the low end, the middle and the top of the band
  {
    auto F = rlc_mna_family(8, 8, 96, 8);
    check("rlc-mna8x8 (probe 1 kHz)", F, 0, 1, 96);
    check("rlc-mna8x8 (probe 30 MHz)", F, 72, 1, 96);
    check("rlc-mna8x8 (probe 10 GHz)", F, 95, 1, 96);
  }
  printf("%s\n", failures ? "FAIL" : "OK");
  return failures ? 1 : 0;
}
