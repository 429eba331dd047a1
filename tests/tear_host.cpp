// Host build of the GPU torn solve (kronspice_amd/csrc/kh_tear.h, compiled
// for the CPU with KH_HD empty) against the oracle's restatement of
// kron.JacobianTearing (oracle/tear.hpp). Prints one line per case; exit
// status 0 when every check holds, for B systems of one pattern solved in
// structure-of-arrays state over several calls:
//  * the plan (pieces, the piece of each unknown) equals the oracle's;
//  * a system the oracle solves torn is solved torn, bit for bit (solution);
//  * a system the oracle hands to the untorn LU is marked for it, and a
//    non-finite one is refused as non-finite;
//  * the counters (solves, pieces factored, reused, reused within the
//    tolerance, fallbacks, refinement steps) equal the oracle's.
// Systems: the reference's 8-unknown chain torn in the middle, random
// clustered systems (diagonally dominant pieces, couplings across the cut,
// off-diagonal pivots, real and complex), and call sequences that reuse
// every piece, change one, move within the reuse tolerance, make a piece
// singular, and plant a NaN.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "kh_tear.h"
#include "../oracle/tear.hpp"

using namespace kh;
using cplx = std::complex<double>;

namespace {

int g_fail = 0;

template <typename T> struct OracleT;
template <> struct OracleT<double> { using type = double; };
template <> struct OracleT<kcplx> { using type = cplx; };

inline cplx to_o(kcplx v) { return {v.re, v.im}; }
inline double to_o(double v) { return v; }
inline bool same_bits(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }
inline bool same_bits(kcplx a, cplx b) { return same_bits(a.re, b.real()) && same_bits(a.im, b.imag()); }
inline bool same_bits(double a, double b, int) { return same_bits(a, b); }

// One batch: the plan, B systems' state, and an oracle matrix per system.
template <typename T>
struct Batch {
  using O = typename OracleT<T>::type;
  int n;
  int64_t B;
  std::vector<int32_t> ri, ci;
  TearHost host;
  std::vector<T> tv;
  std::vector<double> dv;
  std::vector<int32_t> iv;
  std::vector<int64_t> cnt;
  double tol;
  oracle::Tearing jt;
  std::vector<oracle::TornMat<O>> om;
  Batch(int n_, const std::vector<int32_t>& r, const std::vector<int32_t>& c, int64_t B_,
        const std::vector<std::pair<int32_t, int32_t>>& cuts, double tol_)
      : n(n_), B(B_), ri(r), ci(c), tol(tol_) {
    host = tear_plan(n, ri, ci, cuts);
    tv.assign((size_t)host.v.nT * B, lu_zero_value<T>());
    dv.assign((size_t)host.v.nD * B, 0.0);
    iv.assign((size_t)host.v.nI * B, 0);
    cnt.assign((size_t)kTearCounters * B, 0);
    for (const auto& p : cuts) jt.cuts.push_back({p.first, p.second});
    jt.reuseTol = tol;
    om.resize(B);
    for (int64_t s = 0; s < B; s++) om[s].init(n, &jt);
  }
  TearPlanView view() const {
    TearPlanView v = host.v;
    v.ri = host.ri.data(); v.ci = host.ci.data(); v.pieceOf = host.pieceOf.data(); v.local = host.local.data();
    v.rowIdx = host.rowIdx.data(); v.off = host.off.data(); v.rows = host.rows.data(); v.idx = host.idx.data();
    v.pslot = host.pslot.data(); v.pli = host.pli.data(); v.plj = host.plj.data(); v.ports = host.ports.data();
    v.pc = host.pcs.data();
    return v;
  }
  // vals [nslots][B], rhs [n][B]; returns the number of failed checks
  void solve(const char* name, const std::vector<T>& vals, const std::vector<T>& rhs) {
    const TearPlanView v = view();
    std::vector<T> x((size_t)n * B, lu_zero_value<T>());
    std::vector<int> st(B);
    for (int64_t c = 0; c < B; c++) {
      const TearLane<T> L{v, B, c, vals.data(), rhs.data(), x.data(), tv.data(), dv.data(), iv.data(), cnt.data(),
                          tol};
      st[c] = L.solve(host.refused);
    }
    int bad = 0, torn = 0, fb = 0, nf = 0;
    const oracle::Tearing before = jt;
    for (int64_t c = 0; c < B; c++) {
      auto& m = om[c];
      const int64_t fb0 = jt.fallbacks;
      m.clear();
      for (size_t s = 0; s < ri.size(); s++) m.add(ri[s], ci[s], to_o(vals[s * B + c]));
      std::vector<O> b(n), xo(n);
      for (int i = 0; i < n; i++) b[i] = to_o(rhs[(size_t)i * B + c]);
      const int e = m.solve(b.data(), xo.data());
      const bool ofb = jt.fallbacks > fb0;
      if (e == oracle::NONFINITE && !ofb) {
        nf++;
        if (st[c] != LU_NONFINITE) bad++;
        continue;
      }
      if (ofb) {
        fb++;
        if (st[c] != TEAR_FALLBACK) bad++;
        continue;
      }
      torn++;
      if (st[c] != LU_OK) { bad++; continue; }
      for (int i = 0; i < n; i++)
        if (!same_bits(x[(size_t)i * B + c], xo[i])) { bad++; break; }
    }
    int64_t sum[kTearCounters] = {};
    for (int k = 0; k < kTearCounters; k++)
      for (int64_t c = 0; c < B; c++) sum[k] += cnt[(size_t)k * B + c];
    const int64_t want[kTearCounters] = {jt.solves, jt.factors, jt.reuses, jt.stale, jt.fallbacks, jt.refines};
    for (int k = 0; k < kTearCounters; k++)
      if (sum[k] != want[k]) bad++;
    (void)before;
    // the plan: pieces and piece of each unknown (when the oracle has one)
    if (!om.empty() && om[0].plan) {
      if ((int)host.pcs.size() != (int)om[0].plan->pieces.size()) bad++;
      for (int i = 0; i < n; i++)
        if (host.pieceOf[i] != om[0].plan->pieceOf[i]) { bad++; break; }
    }
    std::printf("%-34s B=%-4lld pieces=%-3zu torn=%-4d fallback=%-3d nonfinite=%-3d factors=%-5lld reuses=%-5lld "
                "stale=%-4lld refines=%-4lld %s\n",
                name, (long long)B, host.pcs.size(), torn, fb, nf, (long long)sum[TC_FACTORS],
                (long long)sum[TC_REUSES], (long long)sum[TC_STALE], (long long)sum[TC_REFINES],
                bad ? "FAIL" : "ok");
    if (bad) g_fail++;
  }
};

template <typename T> T rnd_val(std::mt19937_64& g, double lo, double hi);
template <> double rnd_val<double>(std::mt19937_64& g, double lo, double hi) {
  return std::uniform_real_distribution<double>(lo, hi)(g);
}
template <> kcplx rnd_val<kcplx>(std::mt19937_64& g, double lo, double hi) {
  std::uniform_real_distribution<double> d(lo, hi);
  const double a = d(g);
  return {a, d(g) * 0.3};
}
template <typename T> T scale(T v, double s);
template <> double scale<double>(double v, double s) { return v * s; }
template <> kcplx scale<kcplx>(kcplx v, double s) { return {v.re * s, v.im * s}; }

// A clustered system: `k` clusters of `m` unknowns, each a chain with a
// dominant diagonal (and, in every third cluster, a zero diagonal that makes
// the pivot leave it), coupled cluster to cluster by symmetric pairs that the
// cut lists. Returns the pattern and the cut pairs.
void clustered(int k, int m, std::vector<int32_t>& ri, std::vector<int32_t>& ci,
               std::vector<std::pair<int32_t, int32_t>>& cuts, std::mt19937_64& g) {
  const int n = k * m;
  auto add = [&](int i, int j) {
    ri.push_back(i);
    ci.push_back(j);
  };
  for (int q = 0; q < k; q++) {
    const int b = q * m;
    for (int i = 0; i < m; i++) {
      add(b + i, b + i);
      if (i > 0) {
        add(b + i, b + i - 1);
        add(b + i - 1, b + i);
      }
      if (i > 1 && (i % 3) == 0) add(b + i, b + i - 2);  // unsymmetric fill
    }
  }
  for (int q = 0; q + 1 < k; q++) {
    std::uniform_int_distribution<int> u(0, m - 1);
    const int a = q * m + u(g), c = (q + 1) * m + u(g);
    add(a, c);
    add(c, a);
    cuts.push_back({a, c});
  }
  // drop duplicate slots (first occurrence kept), as first-Add order would
  std::vector<int32_t> r2, c2;
  std::vector<char> seen((size_t)n * n, 0);
  for (size_t s = 0; s < ri.size(); s++) {
    const size_t key = (size_t)ri[s] * n + ci[s];
    if (seen[key]) continue;
    seen[key] = 1;
    r2.push_back(ri[s]);
    c2.push_back(ci[s]);
  }
  ri.swap(r2);
  ci.swap(c2);
}

template <typename T>
void values(const std::vector<int32_t>& ri, const std::vector<int32_t>& ci, int64_t B, int m, std::mt19937_64& g,
            std::vector<T>& vals) {
  vals.assign(ri.size() * B, lu_zero_value<T>());
  for (size_t s = 0; s < ri.size(); s++)
    for (int64_t c = 0; c < B; c++) {
      T v = rnd_val<T>(g, -1.0, 1.0);
      if (ri[s] == ci[s]) {
        v = rnd_val<T>(g, 4.0, 8.0);
        if ((ri[s] / m) % 3 == 2 && ri[s] % m == 1) v = lu_zero_value<T>();  // pivot leaves the diagonal
      }
      vals[s * B + c] = v;
    }
}

template <typename T>
void run_family(const char* tag, int k, int m, int64_t B, double tol, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<int32_t> ri, ci;
  std::vector<std::pair<int32_t, int32_t>> cuts;
  clustered(k, m, ri, ci, cuts, g);
  const int n = k * m;
  Batch<T> b(n, ri, ci, B, cuts, tol);
  std::vector<T> vals, rhs((size_t)n * B);
  values<T>(ri, ci, B, m, g, vals);
  for (auto& r : rhs) r = rnd_val<T>(g, -1.0, 1.0);
  char name[96];
  std::snprintf(name, sizeof name, "%s first", tag);
  b.solve(name, vals, rhs);
  std::snprintf(name, sizeof name, "%s same values", tag);
  b.solve(name, vals, rhs);
  // one cluster's diagonal changes in every system
  std::vector<T> v2 = vals;
  for (size_t s = 0; s < ri.size(); s++)
    if (ri[s] == ci[s] && ri[s] / m == 1)
      for (int64_t c = 0; c < B; c++) v2[s * B + c] = scale(v2[s * B + c], 1.25);
  std::snprintf(name, sizeof name, "%s one piece changed", tag);
  b.solve(name, v2, rhs);
  // every value moves by a relative 1e-6 (within a reuse tolerance of 0.01)
  std::vector<T> v3 = v2;
  for (auto& v : v3) v = scale(v, 1 + 1e-6);
  std::snprintf(name, sizeof name, "%s small move", tag);
  b.solve(name, v3, rhs);
  // a singular piece in odd systems: cluster 0's first row zeroed
  std::vector<T> v4 = v3;
  for (size_t s = 0; s < ri.size(); s++)
    if (ri[s] == 0)
      for (int64_t c = 1; c < B; c += 2) v4[s * B + c] = lu_zero_value<T>();
  std::snprintf(name, sizeof name, "%s singular piece", tag);
  b.solve(name, v4, rhs);
  // a NaN in every fourth system
  std::vector<T> v5 = v3;
  for (int64_t c = 0; c < B; c += 4) {
    T& v = v5[(ri.size() - 1) * B + c];
    if constexpr (std::is_same<T, double>::value) v = std::nan("");
    else v.im = std::nan("");
  }
  std::snprintf(name, sizeof name, "%s NaN", tag);
  b.solve(name, v5, rhs);
  std::snprintf(name, sizeof name, "%s back to valid", tag);
  b.solve(name, v3, rhs);
  // large spread of magnitudes: the equilibration decides the pivots
  std::vector<T> v6 = v3;
  for (size_t s = 0; s < ri.size(); s++)
    for (int64_t c = 0; c < B; c++) v6[s * B + c] = scale(v6[s * B + c], std::pow(10.0, (double)(ri[s] % 7) - 3));
  std::snprintf(name, sizeof name, "%s row scales 1e-3..1e3", tag);
  b.solve(name, v6, rhs);
}

void chain8() {
  // kron/jacobian_test.go:339-373: 2 on the diagonal, -1 off it, the (3,4)
  // coupling -0.5 and torn
  std::vector<int32_t> ri, ci;
  std::vector<double> v;
  for (int i = 0; i < 8; i++) {
    ri.push_back(i); ci.push_back(i); v.push_back(2);
    if (i > 0 && i != 4) {
      ri.push_back(i); ci.push_back(i - 1); v.push_back(-1);
      ri.push_back(i - 1); ci.push_back(i); v.push_back(-1);
    }
  }
  ri.push_back(4); ci.push_back(3); v.push_back(-0.5);
  ri.push_back(3); ci.push_back(4); v.push_back(-0.5);
  Batch<double> b(8, ri, ci, 1, {{3, 4}}, 0);
  std::vector<double> rhs(8, 0.0);
  rhs[0] = 1;
  b.solve("chain8 torn at (3,4)", v, rhs);
  b.solve("chain8 again (all reused)", v, rhs);
  if (b.host.pcs.size() != 2 || b.jt.fallbacks != 0 || b.jt.reuses != 2) {
    std::printf("chain8: %zu pieces, %lld fallbacks, %lld reuses: FAIL\n", b.host.pcs.size(),
                (long long)b.jt.fallbacks, (long long)b.jt.reuses);
    g_fail++;
  }
}

}  // namespace

int main() {
  chain8();
  run_family<double>("real 4x6", 4, 6, 16, 0.0, 1);
  run_family<double>("real 4x6 reuse 0.01", 4, 6, 16, 0.01, 2);
  run_family<double>("real 9x12", 9, 12, 8, 0.0, 3);
  run_family<double>("real 3x40 (sparse fallback)", 3, 40, 4, 0.0, 4);
  run_family<kcplx>("complex 4x6", 4, 6, 16, 0.0, 5);
  run_family<kcplx>("complex 5x7 reuse 0.01", 5, 7, 8, 0.01, 6);
  std::printf("%s\n", g_fail ? "FAILED" : "all ok");
  return g_fail ? 1 : 0;
}
