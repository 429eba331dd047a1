// kh_group.h — per-circuit (array-of-structures) views of a group's arrays.
//
// One group = the circuits of one topology (own n, element table, patterns,
// arrays). Storage is per circuit and contiguous: circuit c's value slots are
// val[c*nslots ...], its unknowns x[c*n ...], and so on, so one circuit's
// working set shares cache lines and its serial factorization hits L1/L2.
#pragma once
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
};

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

// One circuit's persistent loop: factor + solve the pending system, advance
// the analysis to its next solve (stamping it), until the analysis ends, the
// stamps need a new ordering, or max_solves solves were done.
KH_HD inline void run_circuit(const Group& g, int64_t cc, int32_t max_solves, bool* running, bool* waiting) {
  CtlArgs a;
  LuArgs la;
  circuit_view(g, cc, a, la);
  Ctl ctl(a, 0);
  int32_t* lus = const_cast<int32_t*>(a.lu_status);
  int solves = 0;
  uint64_t cyc_lu = 0, cyc_ctl = 0;
  for (;;) {
    if (ctl.ci[CI_PHASE] == PH_DONE || ctl.ci[CI_WAIT]) break;
    if (ctl.ci[CI_NEED_SOLVE]) {
      if (solves >= max_solves) break;
      uint64_t t0 = KH_CLOCK();
      int st = lu_factor(la, 0);
      if (st == LU_OK) st = lu_solve(la, 0);
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
  ctl.store();
  *running = ctl.ci[CI_PHASE] != PH_DONE;
  *waiting = ctl.ci[CI_WAIT] != 0;
}

}  // namespace kh
