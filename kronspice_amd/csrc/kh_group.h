// kh_group.h — per-circuit (array-of-structures) views of a group's arrays.
//
// One group = the circuits of one topology (own n, element table, patterns,
// arrays). Storage is per circuit and contiguous: circuit c's value slots are
// val[c*nslots ...], its unknowns x[c*n ...], and so on, so one circuit's
// working set shares cache lines and its serial factorization hits L1/L2.
#pragma once
#include <type_traits>

#include "kh_control.h"

// Cycle counter for the per-circuit split of time between factorization and
// the rest of the analysis (reported in kh_stats); 0 where none is provided.
#ifndef KH_CLOCK
#define KH_CLOCK() ((uint64_t)0)
#endif

namespace kh {

struct Group {
  CtlArgs ctl;  // base pointers of the group's per-circuit arrays; ctl.B = circuits
  LuArgs lu;
  int64_t block0;  // first workgroup of the group
  int32_t gid;
  int32_t cpb;     // circuits per workgroup
  int32_t npar, nstate;
  int32_t lds_bytes;  // > 0: each circuit's staged arrays take this much LDS
  int32_t stage_all;  // every array the factorization touches is staged: it then addresses LDS directly
  uint64_t stage_mask;  // bit i: the i-th array of for_each_array is staged
};

// Visits the pointer field of every per-circuit array of a group (null until
// the array is allocated; the LU workspace is allocated with the first pattern):
// f(field, bytes of one circuit's slice, copy in, copy out, rank). Circuit c's slice
// starts at field + c * bytes. "Copy in" marks what a circuit reads before
// writing within one launch (its state between launches), "copy out" what it
// changes that outlives the launch; the rest is scratch of one factorization
// or one convergence test. Pattern ids and per-circuit transient
// specifications are read-only; parameters change only where a DC sweep
// overrides a source's value. The DC sweep's per-point outputs are not here:
// they are written once per point, straight to device memory. The rank
// orders the arrays for partial staging by how often the state machine and
// the factorization touch them: 0 control words, 1 the unknowns and
// right-hand side, 2 the factorization's vectors and column pointers, 3 the
// stamped values, 4 the factors, 5 the reach lists and the depth-first
// search's stacks (the rest of the factorization's working set), 6 device
// state, 7 parameters, 8 vectors one analysis step uses once.
template <typename G, typename F>
KH_HD inline void for_each_array(G& g, F&& f) {
  const int64_t n = g.ctl.ckt.n, ns = g.ctl.ckt.nslots;
  const int64_t capL = g.lu.capL, capU = g.lu.capU, capR = g.lu.capR;
  auto& a = g.ctl;
  auto& l = g.lu;
  f(a.par, g.npar * 8, true, true, 7);
  f(a.st, g.nstate * 8, true, true, 6);
  f(a.val, ns * 8, true, true, 3);
  f(a.rhs, n * 8, true, true, 1);
  f(a.x, n * 8, true, true, 1);
  f(a.xold, n * 8, true, true, 1);
  f(a.xsave, n * 8, true, true, 8);
  f(a.ptprev, n * 8, true, true, 8);
  f(a.xacc, n * 8, true, true, 8);
  f(a.absbuf, n * 8, false, false, 8);
  f(a.cd, (int64_t)NCD * 8, true, true, 0);
  if (a.lane_spec) f(a.lane_spec, 4 * 8, true, false, 8);
  if (a.nodeset) f(a.nodeset, (int64_t)(a.ckt.numNodes - 1) * 8, true, false, 8);
  if (a.ckt.dense) {  // lu_dense's workspace: matrix, column maxima, y, pivot rows
    f(l.lx, n * n * 8, false, false, 4);
    f(l.udiag, n * 8, false, false, 2);
    f(l.w, n * 8, false, false, 2);
    f(a.ci, (int64_t)NCI * 4, true, true, 0);
    f(a.lu_status, 4, true, true, 0);
    f(l.prow, n * 4, false, false, 2);
    f(l.err_step, 4, true, true, 0);
    f(a.present, ns, true, true, 3);
    return;
  }
  f(l.xw, n * 8, false, false, 2);
  f(l.w, n * 8, false, false, 2);
  f(l.udiag, n * 8, false, false, 2);
  f(l.lx, capL * 8, false, false, 4);
  f(l.ux, capU * 8, false, false, 4);
  f(a.ci, (int64_t)NCI * 4, true, true, 0);
  f(a.lu_status, 4, true, true, 0);
  f(l.pid, 4, true, false, 0);
  f(l.li, capL * 4, false, false, 4);
  f(l.ui, capU * 4, false, false, 4);
  f(l.lp, (n + 1) * 4, false, false, 2);
  f(l.up, (n + 1) * 4, false, false, 2);
  f(l.pinv, n * 4, false, false, 2);
  f(l.prow, n * 4, true, true, 2);
  f(l.xi, n * 4, false, false, 5);
  f(l.pstack, n * 4, false, false, 5);
  f(l.mark, n * 4, false, false, 5);
  f(l.err_step, 4, true, true, 0);
  f(l.rch, capR * 4, true, true, 5);
  f(l.rp, (n + 1) * 4, true, true, 2);
  f(l.dk, n * 4, true, true, 2);
  f(l.cpid, 4, true, true, 0);
  f(a.present, ns, true, true, 3);
}

KH_HD inline int64_t round8(int64_t b) { return (b + 7) & ~(int64_t)7; }

// Bytes of LDS one circuit of the group needs to stage every array (each
// array 8-byte aligned).
KH_HD inline int64_t lds_need(const Group& g) {
  int64_t total = 0;
  for_each_array(g, [&](const auto&, int64_t bytes, bool, bool, int) { total += round8(bytes); });
  return total;
}

// Chooses the arrays a circuit stages within `budget` bytes of LDS: all of
// them when they fit, otherwise the allocated arrays rank by rank (within a
// rank in for_each_array order), each taken when it still fits. Sets
// stage_mask, lds_bytes (0: nothing worth staging; the count of 8-byte words
// is made odd so that the circuits of one wavefront start in different LDS
// banks) and stage_all: 1 when every array of the sparse factorization's
// arguments (LuArgs but the pattern pool) is staged.
inline void plan_staging(Group& g, int64_t budget) {
  g.stage_mask = 0;
  g.stage_all = 0;
  g.lds_bytes = 0;
  int nar = 0;
  for_each_array(g, [&](const auto&, int64_t, bool, bool, int) { nar++; });
  if (nar > 64) return;
  budget -= 8;  // room for the odd-word pad
  int64_t used = 0;
  const int64_t full = lds_need(g);
  if (full <= budget) {
    g.stage_mask = nar == 64 ? ~0ull : ((1ull << nar) - 1);
    g.stage_all = 1;
    used = full;
  } else {
    for (int rank = 0; rank <= 9; rank++) {
      int i = 0;
      for_each_array(g, [&](const auto& p, int64_t bytes, bool, bool, int r) {
        if (r == rank && p && used + round8(bytes) <= budget) {
          g.stage_mask |= 1ull << i;
          used += round8(bytes);
        }
        i++;
      });
    }
  }
  if (used < 64) {  // nothing but a few control words: not worth the copies
    g.stage_mask = 0;
    g.stage_all = 0;
    return;
  }
  if (((used / 8) & 1) == 0) used += 8;
  g.lds_bytes = (int32_t)used;
  if (!g.stage_all && !g.ctl.ckt.dense) {
    // the factorization's arrays are for_each_array fields, or alias one (val, rhs, x)
    const void* lu[] = {g.lu.pid, g.lu.val, g.lu.rhs, g.lu.x, g.lu.xw, g.lu.w, g.lu.udiag, g.lu.lx, g.lu.ux,
                        g.lu.li, g.lu.ui, g.lu.lp, g.lu.up, g.lu.pinv, g.lu.prow, g.lu.xi, g.lu.pstack,
                        g.lu.mark, g.lu.err_step, g.lu.rch, g.lu.rp, g.lu.dk, g.lu.cpid};
    bool all = true;
    for (const void* q : lu) {
      bool found = false;
      int i = 0;
      for_each_array(g, [&](const auto& p, int64_t, bool, bool, int) {
        if (p && (const void*)p == q && ((g.stage_mask >> i) & 1)) found = true;
        i++;
      });
      if (!q || !found) all = false;
    }
    g.stage_all = all ? 1 : 0;
  }
}

// Visits the arrays of `mask`: f(field, bytes of one circuit's slice, copy
// in, copy out, offset in the circuit's LDS slice).
template <typename G, typename F>
KH_HD inline void for_each_staged(G& g, uint64_t mask, F&& f) {
  int64_t off = 0;
  int i = 0;
  for_each_array(g, [&](auto& p, int64_t bytes, bool in, bool out, int) {
    if ((mask >> i) & 1) {
      f(p, bytes, in, out, off);
      off += round8(bytes);
    }
    i++;
  });
}

KH_HD inline int64_t sweep_points(const CtlParams& p) {
  return p.analysis == AN_DC ? (int64_t)p.swCount[0] * (p.swN > 1 ? p.swCount[1] : 1) : 0;
}

// The group as circuit cc sees it once its staged arrays are copied to `base`
// (the staged arrays in for_each_array order, each 8-byte aligned); the
// arrays left in device memory are circuit cc's slices there: run_circuit(view,
// 0, ...) then works on the copies and the slices.
KH_HD inline Group lds_view(const Group& g, char* base, int64_t cc) {
  Group v = g;
  const int64_t pts = sweep_points(g.ctl.p);
  if (v.ctl.sw_x) v.ctl.sw_x += cc * pts * g.ctl.ckt.n;
  if (v.ctl.sw_it) v.ctl.sw_it += cc * pts;
  if (v.ctl.sw_i) v.ctl.sw_i += cc * pts * g.ctl.ckt.nel;
  int64_t off = 0;
  int i = 0;
  for_each_array(v, [&](auto& p, int64_t bytes, bool, bool, int) {
    using Ptr = std::remove_reference_t<decltype(p)>;
    if ((g.stage_mask >> i) & 1) {
      p = (Ptr)(base + off);
      off += round8(bytes);
    } else if (p) {
      p = (Ptr)((const char*)p + cc * bytes);
    }
    i++;
  });
  v.lu.val = v.ctl.val;
  v.lu.rhs = v.ctl.rhs;
  v.lu.x = v.ctl.x;
  v.ctl.B = 1;
  v.lu.B = 1;
  return v;
}

// The argument blocks of circuit c: base pointers advanced to the circuit and
// unit stride, so the per-circuit device code (kh_control.h, kh_lu.h) runs
// unchanged on contiguous storage.
KH_HD inline void circuit_view(const Group& g, int64_t c, CtlArgs& a, LuArgs& l) {
  const int64_t n = g.ctl.ckt.n, ns = g.ctl.ckt.nslots;
  a = g.ctl;
  l = g.lu;
  a.B = 1;
  a.par += c * g.npar;
  a.st += c * g.nstate;
  a.val += c * ns;
  a.present += c * ns;
  a.rhs += c * n; a.x += c * n; a.xold += c * n; a.xsave += c * n; a.ptprev += c * n; a.xacc += c * n;
  a.absbuf += c * n;
  a.ci += c * NCI;
  a.cd += c * NCD;
  a.lu_status += c;
  if (a.lane_spec) a.lane_spec += c * 4;
  if (a.nodeset) a.nodeset += c * (a.ckt.numNodes - 1);
  const int64_t pts = sweep_points(a.p);
  if (a.sw_x) a.sw_x += c * pts * n;
  if (a.sw_it) a.sw_it += c * pts;
  if (a.sw_i) a.sw_i += c * pts * a.ckt.nel;
  l.B = 1;
  l.pid += c;
  l.val = a.val; l.rhs = a.rhs; l.x = a.x;
  l.xw += c * n; l.w += c * n; l.udiag += c * n;
  l.lx += c * l.capL; l.li += c * l.capL;
  l.ux += c * l.capU; l.ui += c * l.capU;
  l.lp += c * (n + 1); l.up += c * (n + 1);
  l.pinv += c * n; l.prow += c * n; l.xi += c * n; l.pstack += c * n; l.mark += c * n;
  l.err_step += c;
  l.rch += c * l.capR; l.rp += c * (n + 1); l.dk += c * n; l.cpid += c;
}

// The LU arguments of a circuit staged in LDS, typed as LDS pointers (the
// pattern pool stays in device memory).
KH_HD inline LuArgsT<double, AsLds> lds_lu(const LuArgs& l) {
  using A = LuArgsT<double, AsLds>;
  A r;
  r.n = l.n; r.nslots = l.nslots; r.B = l.B; r.capL = l.capL; r.capU = l.capU; r.capR = l.capR;
  r.pats = l.pats; r.pool = l.pool;
  r.pid = (A::P<const int32_t>)l.pid;
  r.val = (A::P<const double>)l.val; r.rhs = (A::P<const double>)l.rhs;
  r.x = (A::P<double>)l.x; r.xw = (A::P<double>)l.xw; r.w = (A::P<double>)l.w;
  r.udiag = (A::P<double>)l.udiag; r.lx = (A::P<double>)l.lx; r.ux = (A::P<double>)l.ux;
  r.li = (A::P<int32_t>)l.li; r.ui = (A::P<int32_t>)l.ui; r.lp = (A::P<int32_t>)l.lp; r.up = (A::P<int32_t>)l.up;
  r.pinv = (A::P<int32_t>)l.pinv; r.prow = (A::P<int32_t>)l.prow; r.xi = (A::P<int32_t>)l.xi;
  r.pstack = (A::P<int32_t>)l.pstack; r.mark = (A::P<int32_t>)l.mark; r.err_step = (A::P<int32_t>)l.err_step;
  r.rch = (A::P<int32_t>)l.rch; r.rp = (A::P<int32_t>)l.rp; r.dk = (A::P<int32_t>)l.dk;
  r.cpid = (A::P<int32_t>)l.cpid;
  return r;
}

inline namespace KH_CAT(kinds_, KH_KINDS) {

// One circuit's persistent loop: factor + solve the pending system, advance
// the analysis to its next solve (stamping it), until the analysis ends, the
// stamps need a new ordering or the factors more capacity, or max_solves
// solves were done.
// Lds: the group's arrays are staged in LDS (the view of lds_view); the
// FP64 sparse LU then addresses them as LDS.
template <bool Lds = false>
KH_HD inline void run_circuit(const Group& g, int64_t cc, int32_t max_solves, bool* running, bool* waiting) {
  CtlArgs a;
  LuArgs la;
  circuit_view(g, cc, a, la);
  Ctl ctl(a);
  int32_t* lus = const_cast<int32_t*>(a.lu_status);
  int solves = 0;
  uint64_t cyc_lu = 0, cyc_ctl = 0;
  for (;;) {
    if (ctl.ci[CI_PHASE] == PH_DONE || ctl.ci[CI_WAIT]) break;
    if (ctl.ci[CI_NEED_SOLVE]) {
      if (solves >= max_solves) break;
      uint64_t t0 = KH_CLOCK();
      int st;
      if (a.ckt.dense) {
        st = lu_dense<double, true>(la, 0, a.ckt.slot_i, a.ckt.slot_j);
      } else {
        st = LU_UNSURE;
        if (a.p.precision) {
          // FP32 factors + FP64 refinement; a system it cannot certify is
          // redone in FP64 below, so results follow the FP64 path.
          st = lu_factor_mixed<double, true>(la, 0);
          if (st == LU_OK) {
            int passes = 0;
            st = lu_solve_mixed<double, true>(la, 0, &passes);
            if (st == LU_OK) ctl.cd[CD_REFINE] += (double)passes;
          }
          if (st == LU_UNSURE) ctl.cd[CD_REDO] += 1.0;
        }
        if (st == LU_UNSURE) {
          if constexpr (Lds) {
            const LuArgsT<double, AsLds> ll = lds_lu(la);
            st = lu_factor<double, true>(ll, 0);
            if (st == LU_OK) st = lu_solve<double, true>(ll, 0);
          } else {
            st = lu_factor<double, true>(la, 0);
            if (st == LU_OK) st = lu_solve<double, true>(la, 0);
          }
        }
        if (st == LU_CAPACITY) {
          // L/U outgrew the group's capacity (pivoting away from the diagonal
          // fills more than the structural estimate): the host enlarges it and
          // the factorization is redone in the next launch.
          ctl.ci[CI_WAIT] = WAIT_CAPACITY;
          break;
        }
      }
      lus[0] = st;
      solves++;
      cyc_lu += KH_CLOCK() - t0;
    } else if (ctl.ci[CI_PHASE] != PH_INIT) {
      break;  // nothing pending (every path ends in a solve request or finish)
    }
    uint64_t t1 = KH_CLOCK();
    ctl.step();
    cyc_ctl += KH_CLOCK() - t1;
  }
  ctl.cd[CD_CYC_LU] += (double)cyc_lu;
  ctl.cd[CD_CYC_CTL] += (double)cyc_ctl;
  *running = ctl.ci[CI_PHASE] != PH_DONE;
  *waiting = ctl.ci[CI_WAIT] != 0;
}

}  // inline namespace kinds_<KH_KINDS>
}  // namespace kh
