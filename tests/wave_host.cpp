// CPU check of the wavefront-per-circuit kernel (kronspice_amd/csrc/kh_wave.h)
// against the lane-per-circuit kernel's code (kh_group.h run_circuit): both
// device programs compiled for the host, the wavefront's 64 lanes emulated by
// 64 fibers (tests/host_shim/hip/hip_runtime.h), each circuit of a KHB1 batch
// file (kronspice_amd/workloads.py write_batch_file) run to completion by both
// with the host's pattern analysis between launches, as the library's batch
// runner does. Prints one JSON line; exits 1 unless every circuit's status,
// iteration and step counts, solution and ledger are bit-identical.
//
//   wave_host <batch file> [circuits per group] [backend] [max solves per launch]
//
// Test infrastructure only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#define KH_CLOCK() ((uint64_t)0)
#include "kh_symbolic.h"
#include "kh_wave.h"

using namespace kh;

namespace {

struct HostGroup {
  int32_t num_nodes = 0, nel = 0, nparams = 0, B = 0, has_lane = 0;
  std::vector<kh_element> el;
  std::vector<double> params, lane;
};

bool read_exact(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

// One circuit's arrays (B = 1), as the library allocates them per circuit.
struct One {
  int n = 0, ns = 0;
  std::vector<double> par, st, val, rhs, x, xold, xsave, ptprev, xacc, absbuf, cd, lane;
  std::vector<uint8_t> present;
  std::vector<int32_t> ci, lus, pid;
  int64_t cap = 0;
  std::vector<double> xw, w, udiag, lx, ux;
  std::vector<int32_t> li, ui, lp, up, pinv, prow, xi, pstack, mark, err, rch, rp, dk, cpid;
  void alloc_lu(int64_t c) {
    cap = c;
    xw.assign(n, 0); w.assign(n, 0); udiag.assign(n, 0);
    lx.assign(cap, 0); ux.assign(cap, 0); li.assign(cap, 0); ui.assign(cap, 0);
    lp.assign(n + 1, 0); up.assign(n + 1, 0); pinv.assign(n, 0); prow.assign(n, 0); xi.assign(n, 0);
    pstack.assign(n, 0); mark.assign(n, 0); err.assign(1, 0); rch.assign(2 * cap + n, 0); rp.assign(n + 1, 0);
    dk.assign(n, 0); cpid.assign(1, -1);
  }
};

struct Setup {
  Topology T;
  WavePlan W;
  std::vector<int32_t> lists, rowmajor;
  bool dense = false;
};

CtlParams params_for(bool tran, const double spec3[3]) {
  CtlParams p = {};
  p.relTol = 1e-3; p.absTol = 1e-12; p.vnTol = 1e-6; p.gmin = 1e-12; p.ledgerTol = 1e-9;
  p.itl1 = 100; p.itl4 = 10;
  p.analysis = tran ? AN_TRAN : AN_OP;
  if (tran) {
    p.step = spec3[0]; p.stop = spec3[1]; p.start = 0;
    p.maxStep = spec3[2] > 0 ? spec3[2] : std::min(p.step, p.stop / 50);
    p.chgTol = 1e-14; p.trTol = 7; p.minStep = 1e-12 * p.stop; p.maxSteps = 1000000; p.uic = 0; p.method = TRAP;
  }
  return p;
}

Group make_group(const Setup& S, One& o, const CtlParams& p, PatternPool& pool, bool has_lane) {
  const Topology& T = S.T;
  Group g = {};
  CtlArgs& a = g.ctl;
  a.ckt.n = T.n; a.ckt.nel = T.nel; a.ckt.nslots = T.nslots; a.ckt.numNodes = T.num_nodes;
  a.ckt.et = T.et.data(); a.ckt.isV = T.isV.data(); a.ckt.slot_i = T.slot_i.data(); a.ckt.slot_j = T.slot_j.data();
  a.ckt.diag_slot = T.diag_slot.data();
  a.ckt.nonlin = T.nonlin ? 1 : 0;
  a.ckt.dense = S.dense ? 1 : 0;
  a.ckt.order = S.dense ? S.rowmajor.data() : nullptr;
  a.ckt.n_nl = (int32_t)T.nl_idx.size(); a.ckt.n_rx = (int32_t)T.rx_idx.size(); a.ckt.n_src = (int32_t)T.src_idx.size();
  a.ckt.nl = S.lists.data(); a.ckt.rx = a.ckt.nl + a.ckt.n_nl; a.ckt.src = a.ckt.rx + a.ckt.n_rx;
  a.ckt.nchunks = (int32_t)S.W.ch_el.size() - 1; a.ckt.rec_cap = S.W.rec_cap;
  a.ckt.ch_el = S.W.ch_el.data(); a.ckt.ch_t = S.W.ch_t.data(); a.ckt.tgt = S.W.tgt.data();
  a.ckt.tel_ptr = S.W.tel_ptr.data(); a.ckt.tel = S.W.tel.data(); a.ckt.rec_off = S.W.rec_off.data();
  a.ckt.ro_ptr = S.dense ? S.W.ro_ptr_dense.data() : S.W.ro_ptr_sparse.data();
  a.ckt.ro_slot = S.dense ? S.W.ro_slot_dense.data() : S.W.ro_slot_sparse.data();
  a.p = p;
  a.B = 1;
  a.par = o.par.data(); a.st = o.st.data(); a.val = o.val.data(); a.present = o.present.data();
  a.rhs = o.rhs.data(); a.x = o.x.data(); a.xold = o.xold.data(); a.xsave = o.xsave.data(); a.ptprev = o.ptprev.data();
  a.xacc = o.xacc.data(); a.absbuf = o.absbuf.data(); a.ci = o.ci.data(); a.cd = o.cd.data();
  a.lane_spec = has_lane ? o.lane.data() : nullptr;
  a.lu_status = o.lus.data();
  LuArgs& l = g.lu;
  l.n = T.n; l.nslots = T.nslots; l.B = 1;
  l.capL = l.capU = o.cap; l.capR = 2 * o.cap + T.n;
  l.pats = pool.desc.data(); l.pool = pool.pool.data();
  l.pid = o.pid.data(); l.val = o.val.data(); l.rhs = o.rhs.data(); l.x = o.x.data();
  l.xw = o.xw.data(); l.w = o.w.data(); l.udiag = o.udiag.data(); l.lx = o.lx.data(); l.ux = o.ux.data();
  l.li = o.li.data(); l.ui = o.ui.data(); l.lp = o.lp.data(); l.up = o.up.data(); l.pinv = o.pinv.data();
  l.prow = o.prow.data(); l.xi = o.xi.data(); l.pstack = o.pstack.data(); l.mark = o.mark.data();
  l.err_step = o.err.data(); l.rch = o.rch.data(); l.rp = o.rp.data(); l.dk = o.dk.data(); l.cpid = o.cpid.data();
  g.npar = T.nparams; g.nstate = std::max(T.nstate, 1); g.cpb = 1;
  return g;
}

// Runs one circuit to completion; wave selects the wavefront form.
// Returns launches used, or -1 if it did not finish.
int run(const Setup& S, One& o, const CtlParams& p, bool has_lane, bool wave, int max_solves) {
  PatternPool pool;
  std::vector<char> lds;
  if (wave) lds.assign((size_t)wave_lds_bytes(S.T.n, S.W.rec_cap) + 64, 0);
  for (int launch = 0; launch < 200000; launch++) {
    Group g = make_group(S, o, p, pool, has_lane);
    bool running = false, waiting = false;
    if (wave) {
      struct Ctx {
        const Group* g;
        int max_solves;
        char* lds;
        bool run0, wait0;
      } ctx{&g, max_solves, lds.data(), false, false};
      kh_emu::run(
          [](void* p) {
            Ctx* c = static_cast<Ctx*>(p);
            bool r = false, wt = false;
            run_circuit_wave(*c->g, 0, c->max_solves, c->lds, &r, &wt);
            if (threadIdx.x == 0) { c->run0 = r; c->wait0 = wt; }
          },
          &ctx);
      running = ctx.run0;
      waiting = ctx.wait0;
    } else {
      run_circuit(g, 0, max_solves, &running, &waiting);
    }
    if (!running) return launch + 1;
    const int wt = o.ci[CI_WAIT];
    if (wt == WAIT_PATTERN) {
      std::vector<int32_t> slots;
      for (int s = 0; s < S.T.nslots; s++)
        if (o.present[s]) slots.push_back(s);
      std::string key(reinterpret_cast<const char*>(slots.data()), slots.size() * 4);
      auto f = pool.by_key.find(key);
      const int32_t id = f != pool.by_key.end() ? f->second
                                                : pool.add(analyze_pattern(S.T.n, S.T.slot_i, S.T.slot_j, slots), key);
      o.pid[0] = id;
      const int64_t need = pool.max_fill + S.T.n + 16;
      if (!S.dense && need > o.cap) o.alloc_lu(need);
    } else if (wt == WAIT_CAPACITY) {
      o.alloc_lu(2 * o.cap);
    }
    o.ci[CI_WAIT] = 0;
  }
  return -1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: wave_host <batch file> [circuits per group] [backend] [max solves]\n");
    return 2;
  }
  const int per = argc > 2 ? atoi(argv[2]) : 4;
  const int backend = argc > 3 ? atoi(argv[3]) : 0;
  const int max_solves = argc > 4 ? atoi(argv[4]) : 7;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  char magic[4];
  int32_t ngroups = 0, transient = 0;
  double spec3[3];
  if (!read_exact(f, magic, 4) || !read_exact(f, &ngroups, 4) || !read_exact(f, &transient, 4) ||
      !read_exact(f, spec3, sizeof(spec3)))
    return 2;
  std::vector<HostGroup> hg(ngroups);
  for (auto& g : hg) {
    int32_t h[5];
    if (!read_exact(f, h, sizeof(h))) return 2;
    g.num_nodes = h[0]; g.nel = h[1]; g.nparams = h[2]; g.B = h[3]; g.has_lane = h[4];
    g.el.resize(g.nel);
    g.params.resize((size_t)g.nparams * g.B);
    if (!read_exact(f, g.el.data(), g.el.size() * sizeof(kh_element)) ||
        !read_exact(f, g.params.data(), g.params.size() * 8))
      return 2;
    if (g.has_lane) {
      g.lane.resize((size_t)4 * g.B);
      if (!read_exact(f, g.lane.data(), g.lane.size() * 8)) return 2;
    }
  }
  fclose(f);
  const CtlParams p = params_for(transient != 0, spec3);
  int checked = 0, mismatched = 0, unfinished = 0;
  for (size_t gi = 0; gi < hg.size(); gi++) {
    const HostGroup& H = hg[gi];
    for (int c = 0; c < std::min(per, H.B); c++) {
      std::vector<double> pc(H.nparams);
      for (int k = 0; k < H.nparams; k++) pc[k] = H.params[(size_t)k * H.B + c];
      Setup S;
      if (!S.T.build(H.num_nodes, H.el.data(), H.nel, H.nparams, pc.data())) {
        fprintf(stderr, "group %zu: %s\n", gi, S.T.error.c_str());
        return 2;
      }
      S.W = build_wave_plan(S.T, kWaveRecordLimit);
      S.lists = S.T.nl_idx;
      S.lists.insert(S.lists.end(), S.T.rx_idx.begin(), S.T.rx_idx.end());
      S.lists.insert(S.lists.end(), S.T.src_idx.begin(), S.T.src_idx.end());
      S.rowmajor.resize(S.T.nslots);
      for (int k = 0; k < S.T.nslots; k++) S.rowmajor[k] = k;
      std::sort(S.rowmajor.begin(), S.rowmajor.end(), [&](int32_t a, int32_t b) {
        return S.T.slot_i[a] != S.T.slot_i[b] ? S.T.slot_i[a] < S.T.slot_i[b] : S.T.slot_j[a] < S.T.slot_j[b];
      });
      S.dense = backend == 1 || (backend == 0 && S.T.n < 30);
      One run_o[2];
      for (One& o : run_o) {
        const Topology& T = S.T;
        o.n = T.n; o.ns = T.nslots;
        o.par = pc; o.st.assign(std::max(T.nstate, 1), 0); o.val.assign(T.nslots, 0); o.present.assign(T.nslots, 0);
        o.rhs.assign(T.n, 0); o.x.assign(T.n, 0); o.xold.assign(T.n, 0); o.xsave.assign(T.n, 0);
        o.ptprev.assign(T.n, 0); o.xacc.assign(T.n, 0); o.absbuf.assign(T.n, 0);
        o.ci.assign(NCI, 0); o.cd.assign(NCD, 0); o.lus.assign(1, 0); o.pid.assign(1, 0);
        if (H.has_lane) o.lane = {H.lane[c], H.lane[H.B + c], H.lane[2 * H.B + c], H.lane[3 * H.B + c]};
        o.alloc_lu(S.dense ? (int64_t)T.n * T.n : 16);
      }
      const int l0 = run(S, run_o[0], p, H.has_lane != 0, false, max_solves);
      const int l1 = run(S, run_o[1], p, H.has_lane != 0, true, max_solves);
      checked++;
      if (l0 < 0 || l1 < 0) {
        unfinished++;
        fprintf(stderr, "group %zu circuit %d: unfinished (serial %d, wave %d)\n", gi, c, l0, l1);
        continue;
      }
      const One &a = run_o[0], &b = run_o[1];
      bool same = a.x == b.x;
      for (int k : {CI_STATUS, CI_ITERS, CI_ACCEPTED, CI_REJECTED, CI_SOLVES}) same = same && a.ci[k] == b.ci[k];
      for (int k : {CD_LEDGER, CD_MAXREL, CD_PEAK, CD_WORST}) same = same && a.cd[k] == b.cd[k];
      if (!same) {
        mismatched++;
        double dx = 0;
        for (int i = 0; i < a.n; i++) dx = std::max(dx, std::fabs(a.x[i] - b.x[i]));
        fprintf(stderr,
                "group %zu circuit %d (n %d): status %d/%d iters %d/%d accepted %d/%d rejected %d/%d max|dx| %.3g "
                "ledger %.17g/%.17g\n",
                gi, c, a.n, a.ci[CI_STATUS], b.ci[CI_STATUS], a.ci[CI_ITERS], b.ci[CI_ITERS], a.ci[CI_ACCEPTED],
                b.ci[CI_ACCEPTED], a.ci[CI_REJECTED], b.ci[CI_REJECTED], dx, a.cd[CD_LEDGER], b.cd[CD_LEDGER]);
      }
    }
  }
  printf("{\"checked\": %d, \"mismatched\": %d, \"unfinished\": %d}\n", checked, mismatched, unfinished);
  return (mismatched || unfinished) ? 1 : 0;
}
