// kh_wave.h — the wavefront-per-circuit form of the persistent analysis
// kernel (k_wave).
//
// One wavefront owns one circuit at a time. All 64 lanes run the circuit's
// analysis state machine (CtlT, kh_control.h) in step with identical values,
// so its control flow never diverges; every loop over unknowns, elements,
// value slots or a factor column is spread over the lanes:
//
//  * stamping: each lane runs one element's Load in record mode (kh_devices.h
//    Lane::rslot), then each lane sums the records of one target (a value slot
//    or a right-hand-side row) over the elements that can touch it, in element
//    order and call order — exactly the reference's sequence of Add calls per
//    slot (solver/sparse.go:114-143), so every value is bit-identical;
//  * convergence tests, LTE proposals, breakpoints, accepts: lane-parallel with
//    wavefront reductions (conjunctions and minima do not depend on order);
//  * ledger sums and AbsProduct: terms computed lane-parallel, summed in the
//    reference's order (em/energy.go:25-79, solver/sparse.go:481-490);
//  * Gilbert–Peierls LU (solver/sparse.go:199-401) with the working column,
//    the permuted right-hand side, the pivot map and the column pointers in
//    LDS: per column the scatter, the L-column updates of each reach entry
//    (rows in parallel, reach entries in order), the pivot search (a
//    wavefront argmax with the reference's tie rule) and the L / U writes are
//    lane-parallel; the triangular solves likewise over each column's entries.
//
// Circuits come from a queue (largest first); a launch runs until the queue is
// empty, each circuit up to a solve budget or until the host must analyse a
// new pattern, as in k_persist.
#pragma once
#include <hip/hip_runtime.h>

#include "kh_group.h"

namespace kh {

constexpr int kWaveLanes = 64;
// Record entries of one stamping chunk (12 bytes each in LDS).
constexpr int kWaveRecordLimit = 1024;

// LDS bytes a wavefront needs for a circuit of n unknowns whose stamping
// chunks need rec_cap record entries: records, reduction scratch, and the
// LU's working column, permuted right-hand side, pivot map, pivot rows and
// column pointers.
KH_HD inline int64_t wave_lds_bytes(int64_t n, int64_t rec_cap) {
  return rec_cap * 12 + 64 * 4 + 64 * 8 * 6 + 16 * n + 12 * n + 8 * (n + 1) + 64;
}

#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using LdsP = __attribute__((address_space(3))) T*;
#else
template <class T> using LdsP = T*;
#endif

__device__ inline int wlane() { return (int)threadIdx.x; }
// LDS writes of every lane done before any lane reads (one wavefront: its
// LDS operations complete in order once the counter drains).
__device__ inline void lds_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#else
  __syncthreads();  // host emulation of a wavefront (lanes as threads): a barrier
#endif
}
// Global and LDS writes of every lane visible to every lane of the workgroup
// (one wavefront per workgroup).
__device__ inline void wsync() { __syncthreads(); }

__device__ inline bool wany(bool v) { return __ballot(v) != 0; }
__device__ inline bool wall(bool v) { return __ballot(!v) == 0; }
__device__ inline int wcount(bool v) { return __popcll(__ballot(v)); }
// Lanes below this one whose flag is set.
__device__ inline int wrank(bool v) {
  const unsigned long long m = __ballot(v);
  const unsigned long long below = (wlane() == 0) ? 0ull : (~0ull >> (64 - wlane()));
  return __popcll(m & below);
}
__device__ inline double wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o);
    v = u > v ? u : v;
  }
  return v;
}
// Minimum that never selects NaN (a NaN proposal is never taken, as the
// reference's `if p < h` loops never take one).
__device__ inline double wmin(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o);
    v = u < v ? u : v;
  }
  return v;
}
__device__ inline int wmin_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ inline int64_t wsum_i(int64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// The LU working set of one system held in LDS (T = double, or kcplx for
// the complex seam).
template <class T>
struct WLuT {
  LdsP<T> x;           // [n] working column; the solve's y
  LdsP<T> w;           // [n] permuted right-hand side
  LdsP<int32_t> pinv;  // [n]
  LdsP<int32_t> prow;  // [n]
  LdsP<int32_t> lp;    // [n+1]
  LdsP<int32_t> up;    // [n+1]
};
using WLu = WLuT<double>;

// Carves W out of LDS at p: returns the bytes used (each array 16-aligned).
template <class T>
__device__ inline int64_t wlu_carve(char* p, int n, WLuT<T>& W) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* q = p + off;
    off += (bytes + 15) & ~(int64_t)15;
    return q;
  };
  W.x = (LdsP<T>)(T*)take((int64_t)n * sizeof(T));
  W.w = (LdsP<T>)(T*)take((int64_t)n * sizeof(T));
  W.pinv = (LdsP<int32_t>)(int32_t*)take((int64_t)n * 4);
  W.prow = (LdsP<int32_t>)(int32_t*)take((int64_t)n * 4);
  W.lp = (LdsP<int32_t>)(int32_t*)take((int64_t)(n + 1) * 4);
  W.up = (LdsP<int32_t>)(int32_t*)take((int64_t)(n + 1) * 4);
  return off;
}
// LDS bytes of wlu_carve.
KH_HD inline int64_t wlu_lds_bytes(int64_t n, int64_t tsize) {
  auto r = [](int64_t b) { return (b + 15) & ~(int64_t)15; };
  return 2 * r(n * tsize) + 2 * r(n * 4) + 2 * r((n + 1) * 4);
}

// Wavefront argmax over (|x|, row) with the reference's rule: the larger
// magnitude wins, equal magnitudes go to the lower row (solver/sparse.go:
// 262-268); best < 0 means no candidate.
__device__ inline void warg(double& best, int& piv) {
  for (int o = 32; o > 0; o >>= 1) {
    const double b = __shfl_xor(best, o);
    const int p = __shfl_xor(piv, o);
    if (b > best || (b == best && p >= 0 && (piv < 0 || p < piv))) {
      best = b;
      piv = p;
    }
  }
}

// reach, solver/sparse.go:301-352, run by every lane with identical values
// (the same reads and writes): returns top, the reach is xi[top..n).
template <class T>
__device__ inline int wlu_reach(const LuArgsT<T>& a, const WLuT<T>& W, int k, const int32_t* ap,
                                const int32_t* arow) {
  const int n = a.n;
  const int32_t gen = k + 1;
  int top = n;
  for (int p = ap[k]; p < ap[k + 1]; p++) {
    int32_t r = arow[p];
    if (a.mark[r] == gen) continue;
    int head = 0;
    a.xi[0] = r;
    while (head >= 0) {
      int32_t v = a.xi[head];
      int32_t j = W.pinv[v];
      if (a.mark[v] != gen) {
        a.mark[v] = gen;
        if (j >= 0) a.pstack[head] = W.lp[j];
      }
      bool done = true;
      if (j >= 0) {
        int32_t end = W.lp[j + 1];
        for (int32_t q = a.pstack[head]; q < end; q++) {
          int32_t wv = a.li[q];
          if (a.mark[wv] == gen) continue;
          a.pstack[head] = q + 1;
          head++;
          a.xi[head] = wv;
          done = false;
          break;
        }
      }
      if (done) {
        head--;
        top--;
        a.xi[top] = v;
      }
    }
  }
  return top;
}

// factor, solver/sparse.go:199-293, on one circuit's unit-stride arrays with
// the working set of W; the same reach reuse as lu_factor (kh_lu.h). Returns
// the same status on every lane.
template <class T>
__device__ inline int wlu_factor(const LuArgsT<T>& a, const WLuT<T>& W) {
  const int lane = wlane();
  const int n = a.n;
  const int32_t pidc = a.pid[0];
  const PatDesc pd = a.pats[pidc];
  const int32_t* ap = a.pool + pd.ap;
  const int32_t* arow = a.pool + pd.arow;
  const int32_t* aslot = a.pool + pd.aslot;
  double scale = 0.0;
  bool bad = false;
  for (int s = lane; s < a.nslots; s += kWaveLanes) {
    const T v = a.val[s];
    if (!lu_finite(v)) bad = true;
    const double av = lu_mag(v);
    if (av > scale) scale = av;
  }
  if (wany(bad)) return LU_NONFINITE;
  scale = wmax(scale);
  if (scale == 0) {
    if (lane == 0) a.err_step[0] = -1;
    return LU_SINGULAR;
  }
  bool reuse = a.cpid[0] == pidc;
  for (int i = lane; i < n; i += kWaveLanes) {
    W.pinv[i] = -1;
    W.prow[i] = a.prow[i];  // the previous factorization's pivot rows (reach reuse)
    if (!reuse) a.mark[i] = 0;
  }
  wsync();
  if (lane == 0) a.cpid[0] = -1;  // valid again only once this factorization completes
  int nl = 0, nu = 0, nr = 0;
  for (int k = 0; k < n; k++) {
    if (lane == 0) {
      W.lp[k] = nl;
      W.up[k] = nu;
    }
    const int r0 = nr;
    int r1;
    bool diag_in_reach;
    if (reuse) {
      r1 = a.rp[k + 1];
      diag_in_reach = a.dk[k] != 0;
    } else {
      // The depth-first search is serial: lane 0 runs it, the others wait.
      lds_fence();
      int top = 0;
      if (lane == 0) top = wlu_reach(a, W, k, ap, arow);
      top = __shfl(top, 0);
      wsync();
      r1 = r0 + (n - top);
      if (r1 > a.capR) return LU_CAPACITY;
      for (int p = top + lane; p < n; p += kWaveLanes) a.rch[r0 + p - top] = a.xi[p];
      diag_in_reach = a.mark[k] == k + 1;
      if (lane == 0) {
        a.rp[k + 1] = r1;
        a.dk[k] = diag_in_reach ? 1 : 0;
      }
      wsync();
    }
    nr = r1;
    const int m = r1 - r0;
    for (int p = lane; p < m; p += kWaveLanes) W.x[a.rch[r0 + p]] = lu_zero_value<T>();
    lds_fence();
    double colMax = 0.0;
    for (int p = ap[k] + lane; p < ap[k + 1]; p += kWaveLanes) {
      const T v = a.val[aslot[p]];
      W.x[arow[p]] = v;
      const double av = lu_mag(v);
      if (av > colMax) colMax = av;
    }
    colMax = wmax(colMax);
    lds_fence();
    const double tiny = colMax * (double)n * 1e-14;
    // x = L \ A(:,k): reach entries in order, the rows of each L column in parallel.
    for (int b0 = 0; b0 < m; b0 += kWaveLanes) {
      const int p = b0 + lane;
      int r = 0, j = -1, q0 = 0, len = 0;
      if (p < m) {
        r = a.rch[r0 + p];
        j = W.pinv[r];
        if (j >= 0) {
          q0 = W.lp[j];
          len = W.lp[j + 1] - q0;
        }
      }
      const int cnt = min(kWaveLanes, m - b0);
      for (int t = 0; t < cnt; t++) {
        const int jt = __shfl(j, t);
        if (jt < 0) continue;
        const T xj = W.x[__shfl(r, t)];
        if (lu_zero(xj)) continue;
        const int qa = __shfl(q0, t), ql = __shfl(len, t);
        for (int q = lane; q < ql; q += kWaveLanes) {
          const int i = a.li[qa + q];
          const T prod = a.lx[qa + q] * xj;
          W.x[i] = W.x[i] - prod;
        }
        lds_fence();
      }
    }
    // Pivot among the rows not yet pivotal; U entries in reach order.
    double best = -1.0;
    int piv = -1;
    for (int b0 = 0; b0 < m; b0 += kWaveLanes) {
      const int p = b0 + lane;
      int r = 0, pr = -1;
      if (p < m) {
        r = a.rch[r0 + p];
        pr = W.pinv[r];
      }
      const bool isU = p < m && pr >= 0;
      const int cu = wcount(isU);
      if (nu + cu > a.capU) return LU_CAPACITY;
      const int at = nu + wrank(isU);
      if (isU) {
        a.ui[at] = pr;
        a.ux[at] = W.x[r];
      }
      nu += cu;
      if (p < m && pr < 0) {
        const double av = lu_mag(W.x[r]);
        if (av > best || (av == best && r < piv)) {
          best = av;
          piv = r;
        }
      }
    }
    warg(best, piv);
    if (piv < 0 || best <= tiny) {
      if (lane == 0) a.err_step[0] = k;
      return LU_SINGULAR;
    }
    // KLU diagonal preference (solver/sparse.go:274-277).
    if (W.pinv[k] < 0 && diag_in_reach) {
      const double ak = lu_mag(W.x[k]);
      if (ak >= 1e-3 * best && ak > tiny) piv = k;
    }
    if (reuse && piv != W.prow[k]) {
      // Pivot row differs from the cached factorization: later reaches change.
      reuse = false;
      for (int i = lane; i < n; i += kWaveLanes) a.mark[i] = 0;
      wsync();
    }
    const T pv = W.x[piv];
    lds_fence();
    if (lane == 0) {
      a.udiag[k] = pv;
      W.pinv[piv] = k;
      W.prow[k] = piv;
    }
    lds_fence();
    for (int b0 = 0; b0 < m; b0 += kWaveLanes) {
      const int p = b0 + lane;
      int r = 0;
      bool isL = false;
      if (p < m) {
        r = a.rch[r0 + p];
        isL = W.pinv[r] < 0;
      }
      const int cl = wcount(isL);
      if (nl + cl > a.capL) return LU_CAPACITY;
      const int at = nl + wrank(isL);
      if (isL) {
        a.li[at] = r;
        a.lx[at] = lu_div(W.x[r], pv);
      }
      nl += cl;
    }
    wsync();  // this column of L is read by later columns
  }
  if (lane == 0) {
    W.lp[n] = nl;
    W.up[n] = nu;
  }
  lds_fence();
  for (int i = lane; i <= n; i += kWaveLanes) {
    a.lp[i] = W.lp[i];
    a.up[i] = W.up[i];
    if (i < n) a.prow[i] = W.prow[i];
  }
  if (lane == 0) {
    a.rp[0] = 0;
    a.cpid[0] = pidc;
  }
  wsync();
  return LU_OK;
}

// solve, solver/sparse.go:355-401, with the factors wlu_factor left in W and
// in the circuit's arrays. x is written in original unknown order up to the
// first non-finite entry, as the reference does.
template <class T>
__device__ inline int wlu_solve(const LuArgsT<T>& a, const WLuT<T>& W) {
  const int lane = wlane();
  const int n = a.n;
  const PatDesc pd = a.pats[a.pid[0]];
  const int32_t* q = a.pool + pd.q;
  const int32_t* qinv = a.pool + pd.qinv;
  for (int i = lane; i < n; i += kWaveLanes) W.w[qinv[i]] = a.rhs[i];
  lds_fence();
  // Forward: unit L by columns in pivot order; column k+1's entries are
  // loaded while column k is applied.
  {
    int qa = W.lp[0], ql = W.lp[1] - qa;
    int ni = 0;
    T nv = lu_zero_value<T>();
    if (lane < ql) {
      ni = a.li[qa + lane];
      nv = a.lx[qa + lane];
    }
    for (int k = 0; k < n; k++) {
      const int ci = ni, cqa = qa, cql = ql;
      const T cv = nv;
      if (k + 1 < n) {
        qa = W.lp[k + 1];
        ql = W.lp[k + 2] - qa;
        if (lane < ql) {
          ni = a.li[qa + lane];
          nv = a.lx[qa + lane];
        }
      }
      const T yk = W.w[W.prow[k]];
      if (lane == 0) W.x[k] = yk;
      if (!lu_zero(yk)) {
        if (lane < cql) {
          const T prod = cv * yk;
          W.w[ci] = W.w[ci] - prod;
        }
        for (int p = kWaveLanes + lane; p < cql; p += kWaveLanes) {
          const int i = a.li[cqa + p];
          const T prod = a.lx[cqa + p] * yk;
          W.w[i] = W.w[i] - prod;
        }
      }
      lds_fence();
    }
  }
  // Backward: U by columns, last first.
  {
    int qa = W.up[n - 1], ql = W.up[n] - qa;
    int ni = 0;
    T nv = lu_zero_value<T>(), nd = a.udiag[n - 1];
    if (lane < ql) {
      ni = a.ui[qa + lane];
      nv = a.ux[qa + lane];
    }
    for (int k = n - 1; k >= 0; k--) {
      const int ci = ni, cqa = qa, cql = ql;
      const T cv = nv, cd = nd;
      if (k > 0) {
        qa = W.up[k - 1];
        ql = W.up[k] - qa;
        nd = a.udiag[k - 1];
        if (lane < ql) {
          ni = a.ui[qa + lane];
          nv = a.ux[qa + lane];
        }
      }
      const T zk = lu_div(W.x[k], cd);
      lds_fence();
      if (lane == 0) W.x[k] = zk;
      if (!lu_zero(zk)) {
        if (lane < cql) {
          const T prod = cv * zk;
          W.x[ci] = W.x[ci] - prod;
        }
        for (int p = kWaveLanes + lane; p < cql; p += kWaveLanes) {
          const int i = a.ui[cqa + p];
          const T prod = a.ux[cqa + p] * zk;
          W.x[i] = W.x[i] - prod;
        }
      }
      lds_fence();
    }
  }
  int first_bad = n;
  for (int k = lane; k < n; k += kWaveLanes)
    if (!lu_finite(W.x[k])) first_bad = min(first_bad, k);
  first_bad = wmin_i(first_bad);
  for (int k = lane; k < first_bad; k += kWaveLanes) a.x[q[k]] = W.x[k];
  wsync();
  return first_bad < n ? LU_NONFINITE : LU_OK;
}

inline namespace KH_CAT(kinds_, KH_KINDS) {

// LDS areas of the wavefront form besides the LU's.
struct WArea {
  LdsP<int32_t> rslot;  // [rec_cap] record targets
  LdsP<double> rval;    // [rec_cap] record values
  LdsP<int32_t> rcnt;   // [64] calls recorded per element of the chunk (capped)
  LdsP<double> red;     // [6][64] per-element ledger terms of a chunk
};

// The state machine of one circuit run by a whole wavefront.
struct WCtl : CtlT<WCtl> {
  WArea R;
  KH_HD WCtl(const CtlArgs& a, const WArea& r) : CtlT<WCtl>(a), R(r) {}

  __device__ void vzero(double* v) {
    for (int i = wlane(); i < n; i += kWaveLanes) v[i] = 0;
    wsync();
  }
  __device__ void vcopy(double* d, const double* s) {
    for (int i = wlane(); i < n; i += kWaveLanes) d[i] = s[i];
    wsync();
  }
  __device__ void zero_system() {
    for (int i = wlane(); i < n; i += kWaveLanes) A.rhs[i] = 0;
    for (int s = wlane(); s < A.ckt.nslots; s += kWaveLanes) A.val[s] = 0;
    wsync();
  }

  // Stamping: chunks of consecutive elements, one element per lane, records
  // summed per target in element order (see the file comment).
  __device__ void load_all() {
    const int lane = wlane();
    Lane base = L;
    base.mode = ci[CI_MODE];
    base.method = ci[CI_METHOD];
    base.time = cd[CD_TIME];
    base.H = cd[CD_H];
    base.gmin = cd[CD_GMIN];
    base.srcFact = cd[CD_SRCFACT];
    base.initJct = ci[CI_INITJCT] != 0;
    base.srcStepping = ci[CI_SRCSTEP] != 0;
    const Circ& C = A.ckt;
    const int ns = C.nslots;
    bool grew = false, over = false;
    for (int c = 0; c < C.nchunks; c++) {
      const int k0 = C.ch_el[c], k1 = C.ch_el[c + 1];
      const int k = k0 + lane;
      if (k < k1) {
        const int32_t* e = C.et + (int64_t)k * ET_W;
        const int off = C.rec_off[k];
        Lane l = base;
        l.rslot = (int32_t*)(R.rslot + off);
        l.rval = (double*)(R.rval + off);
        l.rn = 0;
        l.rcap = record_capacity(e[ET_KIND]);
        switch (e[ET_KIND]) {
          case K_R: load_R(l, e); break;
          case K_C: load_C(l, e); break;
          case K_L: load_L(l, e); break;
          case K_V: load_V(l, e); break;
          case K_I: load_I(l, e); break;
          case K_D: if (KH_HAS(K_D)) load_D(l, e); break;
          case K_M: if (KH_HAS(K_M)) load_M(l, e); break;
          case K_Q: if (KH_HAS(K_Q)) load_Q(l, e); break;
          case K_E: if (KH_HAS(K_E)) load_ctl(l, e); break;
          case K_F: if (KH_HAS(K_F)) load_ctl(l, e); break;
          case K_G: if (KH_HAS(K_G)) load_ctl(l, e); break;
          case K_H: if (KH_HAS(K_H)) load_ctl(l, e); break;
          case K_K: if (KH_HAS(K_K)) load_K(l, e); break;
        }
        if (l.rn > l.rcap) over = true;
        R.rcnt[lane] = min(l.rn, l.rcap);
      }
      lds_fence();
      for (int t = C.ch_t[c] + lane; t < C.ch_t[c + 1]; t += kWaveLanes) {
        const int tg = C.tgt[t];
        const bool slot = tg < ns;
        double acc = slot ? A.val[tg] : A.rhs[tg - ns];
        bool found = false;
        for (int qq = C.tel_ptr[t]; qq < C.tel_ptr[t + 1]; qq++) {
          const int kk = C.tel[qq];
          const int off = C.rec_off[kk];
          const int cnt = R.rcnt[kk - k0];
          for (int p = 0; p < cnt; p++)
            if (R.rslot[off + p] == tg) {
              acc = acc + R.rval[off + p];
              found = true;
            }
        }
        if (found) {
          if (slot) {
            A.val[tg] = acc;
            if (!A.present[tg]) {
              A.present[tg] = 1;
              grew = true;
            }
          } else {
            A.rhs[tg - ns] = acc;
          }
        }
      }
      wsync();  // the next chunk reuses the records; its sums read these values
    }
    L.grew = L.grew || wany(grew);
    if (wany(over)) stamp_overflow = true;
  }
  bool stamp_overflow = false;

  __device__ void diag_terms() {
    const double ptG = cd[CD_PTG], g = cd[CD_DIAG_GMIN];
    bool grew = false;
    if (ptG > 0 || g > 0) {
      for (int i = wlane(); i < n; i += kWaveLanes) {
        const int s = diag_slot(i);
        const bool v = A.ckt.isV[i] != 0;
        if (ptG > 0) {
          const double gg = v ? ptG : -ptG;
          A.val[s] = A.val[s] + gg;
          if (v) A.rhs[i] += ptG * A.ptprev[i];
          else A.rhs[i] -= ptG * A.ptprev[i];
          if (!A.present[s]) { A.present[s] = 1; grew = true; }
        }
        if (g > 0 && v) {
          A.val[s] = A.val[s] + g;
          if (!A.present[s]) { A.present[s] = 1; grew = true; }
        }
      }
      wsync();
    }
    L.grew = L.grew || wany(grew);
  }

  __device__ void accept(double t, int mode) {  // Sim.accept, analysis/op.go:324-328
    const Lane Lc = this->L;
    for (int q = wlane(); q < A.ckt.n_rx; q += kWaveLanes) {
      const int32_t* e = A.ckt.et + (int64_t)A.ckt.rx[q] * ET_W;
      switch (e[ET_KIND]) {
        case K_C: accept_C(Lc, e, A.x, t, mode); break;
        case K_L: accept_L(Lc, e, A.x, t, mode); break;
        case K_D: if (KH_HAS(K_D)) accept_D(Lc, e, A.x, t, mode); break;
        case K_M: if (KH_HAS(K_M)) accept_M(Lc, e, A.x, t, mode); break;
        case K_Q: if (KH_HAS(K_Q)) accept_Q(Lc, e, A.x, t, mode); break;
        case K_K: if (KH_HAS(K_K)) accept_K(Lc, e, A.x); break;
      }
    }
    wsync();
  }

  __device__ bool converged() {  // analysis/newton.go:59-82 (side-effect free: a conjunction)
    const Lane Lc = this->L;
    const int32_t* et_ = A.ckt.et;
    bool no = false;
    for (int q = wlane(); q < A.ckt.n_nl; q += kWaveLanes) {
      const int32_t* e = et_ + (int64_t)A.ckt.nl[q] * ET_W;
      if (KH_HAS(K_D) && e[ET_KIND] == K_D && Lc.S(e, 6) != 0) no = true;
      if (KH_HAS(K_M) && e[ET_KIND] == K_M && Lc.S(e, MS_LIM) != 0) no = true;
      if (KH_HAS(K_Q) && e[ET_KIND] == K_Q && Lc.S(e, QS_LIM) != 0) no = true;
    }
    if (wany(no)) return false;
    for (int i = wlane(); i < n; i += kWaveLanes) {
      const double xv = A.x[i], old = A.xold[i];
      double tol = A.ckt.isV[i] ? P.vnTol : P.absTol;
      tol += P.relTol * fmax(fabs(xv), fabs(old));
      if (fabs(xv - old) > tol) no = true;
    }
    if (wany(no)) return false;
    for (int q = wlane(); q < A.ckt.n_nl; q += kWaveLanes) {
      const int32_t* e = et_ + (int64_t)A.ckt.nl[q] * ET_W;
      if (KH_HAS(K_D) && e[ET_KIND] == K_D && !conv_D(Lc, e, A.x, P.relTol, P.absTol)) no = true;
      if (KH_HAS(K_M) && e[ET_KIND] == K_M && !conv_M(Lc, e, A.x, P.relTol, P.absTol)) no = true;
      if (KH_HAS(K_Q) && e[ET_KIND] == K_Q && !conv_Q(Lc, e, A.x, P.relTol, P.absTol)) no = true;
    }
    return !wany(no);
  }

  __device__ double next_breakpoint(double t) {  // analysis/tran.go:233-243
    const Lane Lc = this->L;
    double nb = kInf;
    for (int q = wlane(); q < A.ckt.n_src; q += kWaveLanes) {
      const int32_t* e = A.ckt.et + (int64_t)A.ckt.src[q] * ET_W;
      if ((int)Lc.P(e, 2) == 0) continue;
      const double b = wave_next_breakpoint(Lc, e, t);
      if (b < nb) nb = b;
    }
    return wmin(nb);
  }

  __device__ double truncate(double tn, double h, int order) {
    const Lane Lc = this->L;
    TruncCtx tc;
    tc.time = tn; tc.H = h; tc.method = order;
    tc.relTol = P.relTol; tc.absTol = P.absTol; tc.vnTol = P.vnTol; tc.chgTol = P.chgTol; tc.trTol = P.trTol;
    double prop = kInf;
    for (int q = wlane(); q < A.ckt.n_rx; q += kWaveLanes) {
      const int32_t* e = A.ckt.et + (int64_t)A.ckt.rx[q] * ET_W;
      double pk = kInf;
      switch (e[ET_KIND]) {
        case K_C: pk = trunc_C(Lc, e, A.x, tc); break;
        case K_L: pk = trunc_L(Lc, e, A.x, tc); break;
        case K_D: if (KH_HAS(K_D)) pk = trunc_D(Lc, e, A.x, tc); break;
        case K_M: if (KH_HAS(K_M)) pk = trunc_M(Lc, e, A.x, tc); break;
        case K_Q: if (KH_HAS(K_Q)) pk = trunc_Q(Lc, e, A.x, tc); break;
        default: continue;
      }
      if (pk < prop) prop = pk;
    }
    return wmin(prop);
  }

  // roundoff, analysis/op.go:275-282: each row's AbsProduct in the backend's
  // accumulation order (one lane per row), then the weighted sum in row order.
  __device__ double roundoff() {
    double* const ab = A.absbuf;
    for (int i = wlane(); i < n; i += kWaveLanes) {
      double acc = 0;
      for (int p = A.ckt.ro_ptr[i]; p < A.ckt.ro_ptr[i + 1]; p++) {
        const int s = A.ckt.ro_slot[p];
        if (!A.present[s]) continue;
        acc += fabs(A.val[s]) * fabs(A.x[A.ckt.slot_j[s]]);
      }
      ab[i] = acc;
    }
    wsync();
    double sum = 0;
    int i = 0;
    for (; i + kChunk <= n; i += kChunk) {
      double term[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; u++) term[u] = fabs(A.x[i + u]) * (ab[i + u] + fabs(A.rhs[i + u]));
#pragma unroll
      for (int u = 0; u < kChunk; u++) sum += term[u];
    }
    for (; i < n; i++) sum += fabs(A.x[i]) * (ab[i] + fabs(A.rhs[i]));
    return 64 * 0x1p-52 * sum;
  }

  // Ledger folds (em/energy.go:25-79): per element, in element order, the
  // reference's updates of the running sums, from terms the lanes computed.
  struct Sums {
    double inflow = 0, outflow = 0, diss = 0, stor = 0, sup = 0, scale = 0;
  };
  __device__ static void fold_port(Sums& s, double pw, int ch) {
    if (ch >= 0 && pw != 0) {
      s.scale += fabs(pw);
      if (ch == 2 && pw < 0) s.diss += -pw;
      else if (pw > 0) s.inflow += pw;
      else s.outflow += -pw;
    }
  }
  // Port power and channel of element e (the ExteriorPower rules, as the serial
  // ledger applies them).
  __device__ void port_term(const Lane& Lc, const int32_t* e, const OpPt& o, double* pw, int* ch) const {
    const int kind = e[ET_KIND];
    *pw = 0;
    *ch = -1;
    if (kind == K_R) {
      if (Lc.P(e, 0) > 0) { *pw = -o.power; *ch = 2; }
    } else if (kind == K_V || kind == K_I) {
      *pw = -o.power; *ch = 0;
    } else if (kind == K_D || kind == K_M || kind == K_Q) {
      *pw = -(o.power - o.storage); *ch = 2;
    }
  }

  // Ledger of an operating point: BuildLedger + RelativeResidual
  // (em/energy.go:90-119, output/result.go:100-109).
  __device__ double ledger_op() {
    const Lane Lc = this->L;
    const int lane = wlane();
    Sums s;
    LdsP<double> Tpw = R.red, Tst = R.red + 64, Tms = R.red + 128, Tch = R.red + 192;
    for (int k0 = 0; k0 < A.ckt.nel; k0 += kWaveLanes) {
      const int k = k0 + lane;
      if (k < A.ckt.nel) {
        const int32_t* e = A.ckt.et + (int64_t)k * ET_W;
        const OpPt o = operating(Lc, e, A.x);
        double pw;
        int ch;
        port_term(Lc, e, o, &pw, &ch);
        const int kind = e[ET_KIND];
        const bool rx = kind == K_C || kind == K_L || kind == K_D || kind == K_M || kind == K_Q;
        Tpw[lane] = pw;
        Tch[lane] = ch + (rx ? 16 : 0);  // the channel, and whether the element stores energy
        Tst[lane] = rx ? o.storage : 0.0;
        Tms[lane] = ((kind == K_R && Lc.P(e, 0) < 0) || is_controlled(kind)) ? -o.power : 0.0;
      }
      lds_fence();
      const int cnt = min(kWaveLanes, A.ckt.nel - k0);
      for (int t = 0; t < cnt; t++) {
        const int code = (int)Tch[t];
        const bool rx = code >= 14;
        fold_port(s, Tpw[t], rx ? code - 16 : code);
        if (rx) {
          const double st = Tst[t];
          s.stor += st;
          s.scale += fabs(st);
        }
        const double ms = Tms[t];
        if (ms != 0) { s.sup += ms; s.scale += fabs(ms); }
      }
      lds_fence();
    }
    const double res = s.inflow + s.sup - s.outflow - s.diss - s.stor;
    const double ro = roundoff();
    if (s.scale == 0 || fabs(res) <= fmax(1e-21, ro)) return 0;
    return fabs(res / s.scale);
  }

  // recorder.record's per-point check (analysis/tran.go:291-347): the ports of
  // every element in element order, then the model-supplied power of negative
  // resistors, then the storage of the reactive devices.
  __device__ int record(double t, bool imposed) {
    const Lane Lc = this->L;
    const int lane = wlane();
    Sums s;
    LdsP<double> Tpw = R.red, Tch = R.red + 64;
    for (int k0 = 0; k0 < A.ckt.nel; k0 += kWaveLanes) {
      const int k = k0 + lane;
      if (k < A.ckt.nel) {
        const int32_t* e = A.ckt.et + (int64_t)k * ET_W;
        const int kind = e[ET_KIND];
        double pw = 0;
        int ch = -1;
        if (kind == K_R || kind == K_V || kind == K_I || kind == K_D || kind == K_M || kind == K_Q) {
          const OpPt o = operating(Lc, e, A.x);
          port_term(Lc, e, o, &pw, &ch);
        }
        Tpw[lane] = pw;
        Tch[lane] = ch;
      }
      lds_fence();
      const int cnt = min(kWaveLanes, A.ckt.nel - k0);
      for (int q = 0; q < cnt; q++) fold_port(s, Tpw[q], (int)Tch[q]);
      lds_fence();
    }
    for (int k0 = 0; k0 < A.ckt.nel; k0 += kWaveLanes) {  // model-supplied (negative resistors)
      const int k = k0 + lane;
      double ms = 0;
      if (k < A.ckt.nel) {
        const int32_t* e = A.ckt.et + (int64_t)k * ET_W;
        if ((e[ET_KIND] == K_R && Lc.P(e, 0) < 0) || is_controlled(e[ET_KIND])) ms = -operating(Lc, e, A.x).power;
      }
      if (!wany(ms != 0)) continue;
      Tpw[lane] = ms;
      lds_fence();
      const int cnt = min(kWaveLanes, A.ckt.nel - k0);
      for (int q = 0; q < cnt; q++) {
        const double v = Tpw[q];
        if (v != 0) { s.sup += v; s.scale += fabs(v); }
      }
      lds_fence();
    }
    for (int q0 = 0; q0 < A.ckt.n_rx; q0 += kWaveLanes) {  // storage of reactive devices
      const int q = q0 + lane;
      if (q < A.ckt.n_rx) {
        const int32_t* e = A.ckt.et + (int64_t)A.ckt.rx[q] * ET_W;
        Tpw[lane] = operating(Lc, e, A.x).storage;
      }
      lds_fence();
      const int cnt = min(kWaveLanes, A.ckt.n_rx - q0);
      for (int u = 0; u < cnt; u++) {
        const double st = Tpw[u];
        s.stor += st;
        s.scale += fabs(st);
      }
      lds_fence();
    }
    int err = ST_OK;
    if (s.scale > cd[CD_PEAK]) cd[CD_PEAK] = s.scale;
    const double res = fabs(s.inflow + s.sup - s.outflow - s.diss - s.stor);
    if (!imposed && res > fmax(1e-21, roundoff())) {
      if (res > cd[CD_WORST]) cd[CD_WORST] = res;
      if (res / cd[CD_PEAK] > 1e-3) err = ST_LEDGER;
    }
    (void)t;
    lds_fence();
    return err;
  }

  __device__ void apply_nodeset() {  // analysis/newton.go:221-243
    vzero(A.x);
    for (int i = wlane(); i < A.ckt.numNodes - 1; i += kWaveLanes) A.x[i] = A.nodeset[i];
    wsync();
    const Lane Lc = this->L;
    for (int q = wlane(); q < A.ckt.n_nl; q += kWaveLanes) {
      const int32_t* e = A.ckt.et + (int64_t)A.ckt.nl[q] * ET_W;
      switch (e[ET_KIND]) {
        case K_D: if (KH_HAS(K_D)) accept_D(Lc, e, A.x, 0, MODE_OP); break;
        case K_M: if (KH_HAS(K_M)) accept_M(Lc, e, A.x, 0, MODE_OP); break;
        case K_Q: if (KH_HAS(K_Q)) accept_Q(Lc, e, A.x, 0, MODE_OP); break;
      }
    }
    wsync();
  }
  __device__ bool pt_moved() {
    bool moved = false;
    for (int i = wlane(); i < n; i += kWaveLanes) {
      const double tol = A.ckt.isV[i] ? P.vnTol : P.absTol;
      const double xv = A.x[i];
      if (fabs(xv - A.ptprev[i]) > P.relTol * fabs(xv) + tol) moved = true;
    }
    return wany(moved);
  }
  __device__ bool polish_done() {
    bool done = true;
    for (int i = wlane(); i < n; i += kWaveLanes) {
      const double xv = A.x[i];
      if (fabs(xv - A.xold[i]) > 1e-12 * fmax(1.0, fabs(xv))) done = false;
    }
    return wall(done);
  }
  __device__ void uic_init() {
    for (int k = wlane(); k < A.ckt.nel; k += kWaveLanes) {
      const int32_t* e = el(k);
      if (e[ET_KIND] == K_C) { L.S(e, 3) = L.P(e, 2) != 0 ? L.P(e, 1) : 0.0; L.S(e, 4) = 0; }
      if (e[ET_KIND] == K_L) { L.S(e, 0) = 0; L.S(e, 1) = L.P(e, 2) != 0 ? L.P(e, 1) : 0.0; }
    }
    wsync();
  }
  __device__ void sweep_store(int point) {
    for (int i = wlane(); i < n; i += kWaveLanes) A.sw_x[(int64_t)point * n + i] = A.x[i];
    if (A.sw_i) {
      const Lane Lc = this->L;
      const int nel_ = A.ckt.nel;
      for (int k = wlane(); k < nel_; k += kWaveLanes)
        A.sw_i[(int64_t)point * nel_ + k] = operating_point(Lc, el(k), A.x).current;
    }
    wsync();
  }
};

// One circuit on one wavefront: factor + solve the pending system, advance
// the analysis, until it ends, waits for the host or has used max_solves
// solves. The control words live in LDS meanwhile.
__device__ inline void run_circuit_wave(const Group& g, int64_t cc, int32_t max_solves, char* lds, bool* running,
                                        bool* waiting) {
  const int lane = wlane();
  CtlArgs a;
  LuArgs la;
  circuit_view(g, cc, a, la);
  const int n = a.ckt.n;
  // LDS layout (wave_lds_bytes)
  char* p = lds;
  WArea R;
  const int rc = a.ckt.rec_cap;
  R.rslot = (LdsP<int32_t>)(int32_t*)p; p += (int64_t)rc * 4;
  p = (char*)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  R.rval = (LdsP<double>)(double*)p; p += (int64_t)rc * 8;
  R.rcnt = (LdsP<int32_t>)(int32_t*)p; p += 64 * 4;
  R.red = (LdsP<double>)(double*)p; p += 64 * 8 * 6;
  WLu W;
  W.x = (LdsP<double>)(double*)p; p += (int64_t)n * 8;
  W.w = (LdsP<double>)(double*)p; p += (int64_t)n * 8;
  W.pinv = (LdsP<int32_t>)(int32_t*)p; p += (int64_t)n * 4;
  W.prow = (LdsP<int32_t>)(int32_t*)p; p += (int64_t)n * 4;
  W.lp = (LdsP<int32_t>)(int32_t*)p; p += (int64_t)(n + 1) * 4;
  W.up = (LdsP<int32_t>)(int32_t*)p;
  // Every lane keeps its own copy of the control words and advances it
  // identically; lane 0 writes them back.
  int32_t* gci = a.ci;
  double* gcd = a.cd;
  int32_t pci[NCI];
  double pcd[NCD];
  for (int i = 0; i < NCI; i++) pci[i] = gci[i];
  for (int i = 0; i < NCD; i++) pcd[i] = gcd[i];
  a.ci = pci;
  a.cd = pcd;
  WCtl ctl(a, R);
  int32_t* lus = const_cast<int32_t*>(a.lu_status);
  int solves = 0;
  uint64_t cyc_lu = 0, cyc_ctl = 0;
  for (;;) {
    if (ctl.ci[CI_PHASE] == PH_DONE || ctl.ci[CI_WAIT]) break;
    if (ctl.ci[CI_NEED_SOLVE]) {
      if (solves >= max_solves) break;
      const uint64_t t0 = KH_CLOCK();
      int st;
      if (ctl.stamp_overflow) {
        st = LU_CAPACITY;
      } else if (a.ckt.dense) {
        // the dense backend (n < 30 under Auto): lane 0 runs the serial
        // factorization on the circuit's own workspace
        st = 0;
        if (lane == 0) st = lu_dense<double, true>(la, 0, a.ckt.slot_i, a.ckt.slot_j);
        st = __shfl(st, 0);
        wsync();
      } else {
        st = wlu_factor(la, W);
        if (st == LU_OK) st = wlu_solve(la, W);
        if (st == LU_CAPACITY) {
          // L/U outgrew the group's capacity: the host enlarges it and the
          // factorization is redone in the next launch.
          ctl.ci[CI_WAIT] = WAIT_CAPACITY;
          break;
        }
      }
      if (lane == 0) lus[0] = st;
      wsync();
      solves++;
      cyc_lu += KH_CLOCK() - t0;
    } else if (ctl.ci[CI_PHASE] != PH_INIT) {
      break;
    }
    const uint64_t t1 = KH_CLOCK();
    ctl.step();
    cyc_ctl += KH_CLOCK() - t1;
  }
  ctl.cd[CD_CYC_LU] += (double)cyc_lu;
  ctl.cd[CD_CYC_CTL] += (double)cyc_ctl;
  *running = ctl.ci[CI_PHASE] != PH_DONE;
  *waiting = ctl.ci[CI_WAIT] != 0;
  if (lane == 0) {
    for (int i = 0; i < NCI; i++) gci[i] = pci[i];
    for (int i = 0; i < NCD; i++) gcd[i] = pcd[i];
  }
  wsync();
}

}  // inline namespace kinds_<KH_KINDS>

// Launches k_wave compiled for the kind set persist_kinds(kinds) on the null
// stream: grid wavefronts take circuits (group gw[i], circuit cw[i]) of the
// work list in order from the counter *next until it passes nwork.
hipError_t kh_launch_wave(unsigned kinds, unsigned grid, size_t lds_bytes, const Group* G, const int32_t* gw,
                          const int32_t* cw, int32_t nwork, int32_t* next, int32_t max_solves, int32_t* counters);

}  // namespace kh
