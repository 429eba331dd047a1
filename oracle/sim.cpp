// ORACLE — test infrastructure only (see sparse.hpp).
//
// CPU restatement of kronspice's analysis loop: Sim.solve / newton / dcSolve
// and its fallbacks (analysis/op.go, analysis/newton.go), the operating point
// with its ledger (analysis/op.go:331-404, em/energy.go), the transient loop
// (analysis/tran.go) and the AC sweep (analysis/ac.go). Exposed through a small
// C API for ctypes (tests) and a threaded batch runner (bench.py's CPU
// baseline).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "devices.hpp"
#include "sparse.hpp"
#include "tear.hpp"

namespace oracle {

// acBranch, device/bjt.go:721-734
static void acBranch(CStamper* m, int a, int b, cplx y1, int p1, int n1, cplx y2, int p2, int n2) {
  if (y1 != cplx(0)) {
    m->add(a, p1, y1); m->add(a, n1, -y1); m->add(b, p1, -y1); m->add(b, n1, y1);
  }
  if (y2 != cplx(0)) {
    m->add(a, p2, y2); m->add(a, n2, -y2); m->add(b, p2, -y2); m->add(b, n2, y2);
  }
}

// MOSFET LoadAC, device/mosfet.go:779-814
void Mosfet::loadAC(ACContext& ctx) {
  double md, cd, gm_, gds_, gmbs_, v_on, v_dsat;
  channel(acc.vgs, acc.vds, acc.vbs, &md, &cd, &gm_, &gds_, &gmbs_, &v_on, &v_dsat);
  if (md > 0) { ac.gch[0] = gm_; ac.gch[1] = gds_; ac.gch[2] = gmbs_; }
  else { ac.gch[0] = -gm_; ac.gch[1] = gm_ + gds_ + gmbs_; ac.gch[2] = -gmbs_; }
  double tmp;
  bulkDiode(acc.vbs, iss, &tmp, &ac.gbs);
  bulkDiode(acc.vbd, isd, &tmp, &ac.gbd);
  double mc[3];
  meyerCaps(acc.vgs, acc.vgd, v_on, v_dsat, md, mc);
  double qq, cbs_, cbd_;
  bulkCharge(acc.vbs, czbs, czbssw, &qq, &cbs_);
  bulkCharge(acc.vbd, czbd, czbdsw, &qq, &cbd_);
  ac.cap[0] = mc[0] + cgso; ac.cap[1] = mc[1] + cgdo; ac.cap[2] = mc[2] + cgbo; ac.cap[3] = cbs_; ac.cap[4] = cbd_;
  ac.omega = ctx.omega;
  int ctrlP[3] = {g, dp, b}, ctrlN[3] = {sp, sp, sp};
  for (int k = 0; k < 3; k++) acBranch(ctx.M, dp, sp, cplx(ac.gch[k], 0), ctrlP[k], ctrlN[k], 0, 0, 0);
  stampY(ctx.M, b, sp, cplx(ac.gbs, 0));
  stampY(ctx.M, b, dp, cplx(ac.gbd, 0));
  int ends[5][2] = {{g, sp}, {g, dp}, {g, b}, {b, sp}, {b, dp}};
  for (int k = 0; k < 5; k++) stampY(ctx.M, ends[k][0], ends[k][1], cplx(0, ctx.omega * ac.cap[k]));
  if (dp != u[0]) stampY(ctx.M, u[0], dp, cplx(gdpr, 0));
  if (sp != u[1]) stampY(ctx.M, u[1], sp, cplx(gspr, 0));
}

// Options.Solver for the analyses the tests run through the C entry points:
// when set, every real and complex system an analysis makes is a torn one of
// this JacobianTearing (analysis/op.go:228-241 useSolver, :73-78 cmatrix).
static Tearing* g_tear = nullptr;

// Real system: Sparse or Dense behind one stamper (solver/matrix.go:23-36,
// :77-93), or the factory's torn matrix.
struct RealMatrix : Stamper {
  bool sparse = true;
  SparseLU<double> s;
  Dense d;
  std::unique_ptr<TornMat<double>> t;
  int n = 0;
  int init(int n_, bool sp) {
    n = n_;
    sparse = sp;
    if (g_tear) {
      t = std::make_unique<TornMat<double>>();
      t->init(n_, g_tear);
      return OK;
    }
    return sp ? s.init(n_) : d.init(n_);
  }
  void add(int i, int j, double v) override { if (t) t->add(i, j, v); else if (sparse) s.add(i, j, v); else d.add(i, j, v); }
  void clear() { if (t) t->clear(); else if (sparse) s.clear(); else d.clear(); }
  int solve(const double* b, double* x) { return t ? t->solve(b, x) : sparse ? s.solve(b, x) : d.solve(b, x); }
  void absProduct(const double* x, double* y) const {
    if (t) t->absProduct(x, y);
    else if (sparse) s.absProduct(x, y);
    else d.absProduct(x, y);
  }
};

// Complex system: CSparse or CDense (solver/matrix.go:95-101), or the
// factory's torn complex matrix.
struct ComplexMatrix : CStamper {
  bool sparse = true;
  SparseLU<cplx> s;
  CDense d;
  std::unique_ptr<TornMat<cplx>> t;
  int init(int n_, bool sp) {
    sparse = sp;
    if (g_tear) {
      t = std::make_unique<TornMat<cplx>>();
      t->init(n_, g_tear);
      return OK;
    }
    return sp ? s.init(n_) : d.init(n_);
  }
  void add(int i, int j, cplx v) override { if (t) t->add(i, j, v); else if (sparse) s.add(i, j, v); else d.add(i, j, v); }
  void clear() { if (t) t->clear(); else if (sparse) s.clear(); else d.clear(); }
  int solve(const cplx* b, cplx* x) { return t ? t->solve(b, x) : sparse ? s.solve(b, x) : d.solve(b, x); }
  void absProduct(const cplx* x, double* y) const {
    if (t) t->absProduct(x, y);
    else if (sparse) s.absProduct(x, y);
    else d.absProduct(x, y);
  }
};

// Backend.Resolve, solver/matrix.go:77-85: 0 auto, 1 dense, 2 sparse.
static bool resolve_sparse(int backend, int n) { return backend == 2 || (backend == 0 && n >= 30); }

struct Options {  // analysis.Options with defaults(), analysis/op.go:91-110
  double relTol = 1e-3, absTol = 1e-12, vnTol = 1e-6, gmin = 1e-12, ledgerTol = 1e-9;
  int itl1 = 100, itl4 = 10;
  bool noGmin = false, noSource = false, noPT = false, skipEnergy = false;
  int backend = 2;  // 0 auto, 1 dense, 2 sparse
  void from(const kh_options* o) {
    if (!o) return;
    relTol = o->reltol > 0 ? o->reltol : 1e-3;
    absTol = o->abstol > 0 ? o->abstol : 1e-12;
    vnTol = o->vntol > 0 ? o->vntol : 1e-6;
    gmin = o->gmin > 0 ? o->gmin : 1e-12;
    itl1 = o->itl1 > 0 ? o->itl1 : 100;
    itl4 = o->itl4 > 0 ? o->itl4 : 10;
    noGmin = o->no_gmin_stepping != 0;
    noSource = o->no_source_stepping != 0;
    noPT = o->no_pseudo_transient != 0;
    skipEnergy = o->skip_energy_check != 0;
    ledgerTol = o->ledger_tol > 0 ? o->ledger_tol : 1e-9;
  }
};

enum { E_OK = 0, E_SINGULAR = 1, E_NONFINITE = 2, E_BADSIZE = 3, E_NOCONV = 4, E_TSTEP = 5, E_MAXSTEPS = 6,
       E_LEDGER = 8, E_BADSPEC = 9, E_UNSUPPORTED = 10 };

static bool retryable(int e) { return e == E_NOCONV || e == E_NONFINITE || e == E_SINGULAR; }  // newton.go:86-89

// Planted mistakes (tests only): stepping constants of analysis/newton.go that
// a test may alter to show its checks notice the change. 0 = the reference's.
static double g_plant_gmin_cap = 0;     // factor cap of gminStepping (10, newton.go:153)
static double g_plant_src_growth = 0;   // step growth of sourceStepping (1.5, newton.go:205)
static double g_plant_pt_growth = 0;    // step growth of pseudoTransient (2, newton.go:309)

struct Sim {
  int numNodes = 1;
  // Optional trace of the continuation: per Newton call (method, parameter,
  // status), method 0 plain, 1 gmin (diagGmin), 2 source (SrcFact),
  // 3 pseudo-transient (step h).
  std::vector<double>* trace = nullptr;
  double ptH = 0;
  int backendSel = 0;  // Options.Backend the Sim was built with (the AC matrix resolves it too)
  std::vector<std::unique_ptr<Device>> devs;
  std::vector<int> reactive, nonlin, breakers;
  std::vector<char> isV;
  std::vector<double> xOld, x, rhs, absBuf, ptPrev;
  double diagGmin = 0, ptG = 0;
  bool warm = false;
  RealMatrix m;
  LoadContext lc;
  int n = 0;
  int64_t solves = 0;
  int lastErr = 0;

  int build(int nNodes, const kh_element* el, int nel, const double* params, int backend, double step, double stop) {
    numNodes = nNodes;
    SetupContext ctx;
    ctx.numNodes = nNodes;
    for (int e = 0; e < nel; e++) {
      const double* p = params + el[e].poff;
      std::unique_ptr<Device> d;
      switch (el[e].kind) {
        case KH_ELEM_R: { auto r = new Resistor(); r->R = p[0]; r->g = p[1]; d.reset(r); break; }
        case KH_ELEM_C: { auto c = new Capacitor(); c->C = p[0]; c->ic = p[1]; c->hasIC = p[2] != 0; d.reset(c); break; }
        case KH_ELEM_L: { auto l = new Inductor(); l->L = p[0]; l->ic = p[1]; l->hasIC = p[2] != 0; d.reset(l); break; }
        case KH_ELEM_V: { auto v = new VoltageSource(); loadSource(v->src, p); d.reset(v); break; }
        case KH_ELEM_I: { auto v = new CurrentSource(); loadSource(v->src, p); d.reset(v); break; }
        case KH_ELEM_D: {
          auto dd = new Diode();
          dd->IS = p[0]; dd->N = p[1]; dd->RS = p[2]; dd->CJO = p[3]; dd->VJ = p[4]; dd->M = p[5]; dd->TT = p[6];
          dd->BV = p[7]; dd->IBV = p[8]; dd->FC = p[9]; dd->hasBV = p[10] != 0; dd->vte = p[11]; dd->vcrit = p[12];
          dd->xbv = p[13]; dd->grs = p[14]; dd->hasCharge = p[15] != 0;
          d.reset(dd);
          break;
        }
        case KH_ELEM_M: {
          auto q = new Mosfet();
          q->typ = p[0]; q->beta = p[1]; q->vbi = p[2]; q->vto = p[3]; q->GAMMA = p[4]; q->PHI = p[5]; q->LAMBDA = p[6];
          q->gdpr = p[7]; q->gspr = p[8]; q->isd = p[9]; q->iss = p[10]; q->vcritD = p[11]; q->vcritS = p[12];
          q->czbd = p[13]; q->czbdsw = p[14]; q->czbs = p[15]; q->czbssw = p[16]; q->PB = p[17]; q->MJ = p[18];
          q->MJSW = p[19]; q->FC = p[20];
          for (int k = 0; k < 3; k++) { q->fbt[k] = p[21 + k]; q->fsw[k] = p[24 + k]; }
          q->cgso = p[27]; q->cgdo = p[28]; q->cgbo = p[29]; q->oxideCap = p[30]; q->vt = p[31];
          q->hasJcap = p[32] != 0; q->hasGcap = p[33] != 0;
          MosLevel& lv = q->lv;
          lv.level = (int)p[34]; lv.NFS = p[35]; lv.XJ = p[36]; lv.UCRIT = p[37]; lv.UEXP = p[38]; lv.VMAX = p[39];
          lv.NEFF = p[40]; lv.DELTA = p[41]; lv.THETA = p[42]; lv.ETA = p[43]; lv.KAPPA = p[44]; lv.UO = p[45];
          lv.leff = p[46]; lv.weff = p[47]; lv.cox = p[48]; lv.xd = p[49]; lv.alpha = p[50]; lv.cfs = p[51];
          lv.factor = p[52]; lv.narrow = p[53]; lv.eta3 = p[54]; lv.sqrtPB = p[55]; lv.hasNSUB = p[56] != 0;
          q->LD = p[57];
          q->acc.von = q->vto;  // derive, device/mosfet.go:321
          d.reset(q);
          break;
        }
        case KH_ELEM_Q: {
          auto q = new Bjt();
          double* f[37] = {&q->IS, &q->BF, &q->NF, &q->VAF, &q->IKF, &q->ISE, &q->NE, &q->BR, &q->NR, &q->VAR,
                           &q->IKR, &q->ISC, &q->NC, &q->RB, &q->IRB, &q->RBM, &q->RE, &q->RC, &q->CJE, &q->VJE,
                           &q->MJE, &q->TF, &q->XTF, &q->VTF, &q->ITF, &q->CJC, &q->VJC, &q->MJC, &q->XCJC, &q->TR,
                           &q->CJS, &q->VJS, &q->MJS, &q->FC, &q->ISS, &q->NS, &q->SUBS};
          q->typ = p[0];
          for (int k = 0; k < 37; k++) *f[k] = p[1 + k];
          q->vt = p[38]; q->vcrit = p[39]; q->rbpr = p[40]; q->rbpi = p[41]; q->gcpr = p[42]; q->gepr = p[43];
          q->ovtf = p[44]; q->f1e = p[45]; q->f2e = p[46]; q->f3e = p[47]; q->f1c = p[48]; q->f2c = p[49];
          q->f3c = p[50]; q->hasCharge = p[51] != 0;
          d.reset(q);
          break;
        }
        case KH_ELEM_E:
        case KH_ELEM_F:
        case KH_ELEM_G:
        case KH_ELEM_H: {
          auto c = new Controlled();
          c->gain = p[0];
          if (el[e].kind == KH_ELEM_E || el[e].kind == KH_ELEM_G) {
            c->cnode[0] = el[e].node[2];
            c->cnode[1] = el[e].node[3];
          } else {
            c->ctrl = el[e].node[2];
          }
          d.reset(c);
          break;
        }
        case KH_ELEM_K: {
          auto mu = new Mutual();
          mu->M = p[0];
          mu->l[0] = el[e].node[0];
          mu->l[1] = el[e].node[1];
          d.reset(mu);
          break;
        }
        default:
          return E_UNSUPPORTED;
      }
      d->kind = el[e].kind;
      for (int k = 0; k < 4; k++) d->nodes[k] = el[e].kind == KH_ELEM_K ? 0 : el[e].node[k];
      devs.push_back(std::move(d));
    }
    (void)step; (void)stop;
    for (size_t i = 0; i < devs.size(); i++) {  // newSim, analysis/op.go:160-208
      devs[i]->setup(ctx);
      if (devs[i]->reactive()) reactive.push_back((int)i);
      if (devs[i]->nonlinear()) nonlin.push_back((int)i);
      if (devs[i]->breakpointer()) breakers.push_back((int)i);
    }
    for (auto& d : devs) {  // Link, device/controlled.go:67-77: F and H find their controlling branch
      if (d->kind != KH_ELEM_F && d->kind != KH_ELEM_H) continue;
      auto* c = static_cast<Controlled*>(d.get());
      int b = -1;
      if (c->ctrl >= 0 && c->ctrl < (int)devs.size()) {
        Device* t = devs[c->ctrl].get();
        if (auto* v = dynamic_cast<VoltageSource*>(t)) b = v->branch;
        else if (auto* l = dynamic_cast<Inductor*>(t)) b = l->branch;
        else if (auto* o = dynamic_cast<Controlled*>(t)) b = o->vout() ? o->branch : -1;
      }
      if (b < 0) return E_UNSUPPORTED;
      c->ck = b;
    }
    for (auto& d : devs) {  // Link, device/mutual.go:57-72: the coupled inductors' branches
      if (d->kind != KH_ELEM_K) continue;
      auto* mu = static_cast<Mutual*>(d.get());
      int b[2] = {-1, -1};
      for (int t = 0; t < 2; t++)
        if (mu->l[t] >= 0 && mu->l[t] < (int)devs.size())
          if (auto* l = dynamic_cast<Inductor*>(devs[mu->l[t]].get())) b[t] = l->branch;
      if (b[0] < 0 || b[1] < 0 || b[0] == b[1]) return E_UNSUPPORTED;
      mu->k1 = b[0];
      mu->k2 = b[1];
    }
    n = ctx.size();
    bool sp = resolve_sparse(backend, n);
    backendSel = backend;
    int e = m.init(n, sp);
    if (e != OK) return E_BADSIZE;
    rhs.assign(n, 0); x.assign(n, 0); xOld.assign(n, 0); absBuf.assign(n, 0);
    isV = ctx.voltageUnknowns();
    lc.M = &m;
    lc.RHS = &rhs;
    lc.X = xOld.data();
    lc.nx = n;
    lc.gmin = 1e-12;
    return E_OK;
  }

  // Sim.solve, analysis/op.go:289-322
  int solve() {
    std::copy(x.begin(), x.end(), xOld.begin());
    m.clear();
    std::fill(rhs.begin(), rhs.end(), 0.0);
    for (auto& d : devs) d->load(lc);
    if (ptG > 0) {
      double g = ptG;
      for (int i = 0; i < n; i++) {
        if (isV[i]) { m.add(i, i, g); rhs[i] += g * ptPrev[i]; }
        else { m.add(i, i, -g); rhs[i] -= g * ptPrev[i]; }
      }
    }
    if (diagGmin > 0) {
      for (int i = 0; i < n; i++)
        if (isV[i]) m.add(i, i, diagGmin);
    }
    solves++;
    return m.solve(rhs.data(), x.data());
  }

  void accept(double t, int mode) {  // op.go:324-328
    for (auto& d : devs) d->accept(x.data(), n, t, mode);
  }

  // newton + converged, analysis/newton.go:36-82
  int newton(int maxIter, const Options& o, int* iters) {
    if (!trace) return newton_(maxIter, o, iters);
    const double kind = diagGmin > 0 ? 1 : (lc.sourceStepping ? 2 : (ptG > 0 ? 3 : 0));
    const double par = kind == 1 ? diagGmin : (kind == 2 ? lc.srcFact : (kind == 3 ? ptH : 0));
    const int e = newton_(maxIter, o, iters);
    trace->push_back(kind);
    trace->push_back(par);
    trace->push_back(e);
    return e;
  }
  int newton_(int maxIter, const Options& o, int* iters) {
    lc.gmin = o.gmin;
    for (int it = 0; it < maxIter; it++) {
      int e = solve();
      if (iters) (*iters)++;
      if (e != E_OK) return e;
      if (nonlin.empty()) return E_OK;
      if (lc.initJct) { lc.initJct = false; continue; }
      if (converged(o)) return E_OK;
    }
    return E_NOCONV;
  }
  bool converged(const Options& o) {
    for (int i : nonlin)
      if (devs[i]->limited()) return false;
    for (int i = 0; i < n; i++) {
      double xv = x[i], old = xOld[i];
      double tol = isV[i] ? o.vnTol : o.absTol;
      tol += o.relTol * std::max(std::fabs(xv), std::fabs(old));
      if (std::fabs(xv - old) > tol) return false;
    }
    for (int i : nonlin)
      if (!devs[i]->convTest(x.data(), n, o.relTol, o.absTol)) return false;
    return true;
  }

  // dcSolve, analysis/newton.go:95-128 (no NodeSet here: see warmStart)
  int dcSolve(const Options& o, int* iters) {
    lc.initJct = !warm;
    int err = newton(o.itl1, o, iters);
    if (err == E_OK || nonlin.empty() || !retryable(err)) {
      warm = warm || err == E_OK;
      return err;
    }
    int first = err;
    if (!o.noGmin) {
      if ((err = gminStepping(o, iters)) == E_OK) { warm = true; return E_OK; }
    }
    if (!o.noSource) {
      if ((err = sourceStepping(o, iters)) == E_OK) { warm = true; return E_OK; }
    }
    if (!o.noPT) {
      if ((err = pseudoTransient(o, iters)) == E_OK) { warm = true; return E_OK; }
    }
    (void)err;
    return first;
  }

  // gminStepping, analysis/newton.go:134-176
  int gminStepping(const Options& o, int* iters) {
    std::vector<double> save(n, 0.0);
    std::fill(x.begin(), x.end(), 0.0);
    lc.initJct = true;
    double g = 1e-2, factor = 10.0;
    bool ok = false;
    int ret = E_NOCONV;
    for (int k = 0; k < 200; k++) {
      diagGmin = g;
      int err = newton(o.itl1, o, iters);
      if (err == E_OK) {
        save = x;
        if (g == 0) { ret = E_OK; goto done; }
        ok = true;
        factor = std::min(factor * std::sqrt(factor), g_plant_gmin_cap > 0 ? g_plant_gmin_cap : 10.0);
        if (g / factor < o.gmin) g = 0; else g /= factor;
        continue;
      } else if (!retryable(err)) {
        ret = err;
        goto done;
      }
      if (!ok) { ret = E_NOCONV; goto done; }
      x = save;
      {
        double prev = g * factor;
        if (g == 0) prev = o.gmin;
        factor = std::sqrt(factor);
        if (factor < 1.00005) { ret = E_NOCONV; goto done; }
        g = prev / factor;
      }
    }
    ret = E_NOCONV;
  done:
    diagGmin = 0;
    return ret;
  }

  // sourceStepping, analysis/newton.go:180-214
  int sourceStepping(const Options& o, int* iters) {
    std::vector<double> save(n, 0.0);
    std::fill(x.begin(), x.end(), 0.0);
    lc.sourceStepping = true;
    double done = 0.0, step = 0.1;
    lc.srcFact = 0;
    int ret = E_NOCONV;
    int err = newton(o.itl1, o, iters);
    if (err != E_OK) { ret = err; goto out; }
    save = x;
    for (int k = 0; k < 1000; k++) {
      double f = std::min(done + step, 1.0);
      lc.srcFact = f;
      err = newton(o.itl1, o, iters);
      if (err == E_OK) {
        if (f == 1) { ret = E_OK; goto out; }
        save = x;
        done = f;
        step = std::min(step * (g_plant_src_growth > 0 ? g_plant_src_growth : 1.5), 0.5);
        continue;
      } else if (!retryable(err)) {
        ret = err;
        goto out;
      }
      x = save;
      step /= 2;
      if (step < 1e-3) { ret = E_NOCONV; goto out; }
    }
    ret = E_NOCONV;
  out:
    lc.sourceStepping = false;
    lc.srcFact = 0;
    return ret;
  }

  // pseudoTransient, analysis/newton.go:260-311
  int pseudoTransient(const Options& o, int* iters) {
    if (ptPrev.empty()) ptPrev.assign(n, 0.0);
    std::fill(x.begin(), x.end(), 0.0);
    ptPrev = x;
    lc.initJct = true;
    double h = 1e-6;
    int quiet = 0;
    int ret = E_NOCONV;
    for (int k = 0; k < 2000; k++) {
      ptG = 1 / h;
      ptH = h;
      int err = newton(o.itl1, o, iters);
      if (err != E_OK) {
        if (!retryable(err)) { ret = err; goto out; }
        x = ptPrev;
        h /= 4;
        if (h < 1e-15) { ret = E_NOCONV; goto out; }
        continue;
      }
      bool moved = false;
      for (int i = 0; i < n; i++) {
        double tol = isV[i] ? o.vnTol : o.absTol;
        if (std::fabs(x[i] - ptPrev[i]) > o.relTol * std::fabs(x[i]) + tol) { moved = true; break; }
      }
      ptPrev = x;
      if (!moved) quiet++; else quiet = 0;
      if (quiet >= 3 && h >= 1e6) {
        ptG = 0;
        ret = newton(o.itl1, o, iters);
        goto out;
      }
      h = std::min(h * (g_plant_pt_growth > 0 ? g_plant_pt_growth : 2.0), 1e12);
    }
    ret = E_NOCONV;
  out:
    ptG = 0;
    return ret;
  }

  // applyNodeSet, analysis/newton.go:221-243 (by unknown index: node k -> x[k-1])
  void applyGuess(const double* guess) {
    for (int i = 0; i < n; i++) x[i] = 0;
    for (int i = 0; i < numNodes - 1; i++) x[i] = guess[i];
    for (int i : nonlin) devs[i]->accept(x.data(), n, 0, ModeOP);
  }

  // polish, analysis/op.go:251-271
  int polish(int* iters = nullptr) {
    if (nonlin.empty()) return E_OK;
    for (int k = 0; k < 2; k++) {
      int e = solve();
      if (iters) (*iters)++;
      if (e != E_OK) return e;
      bool done = true;
      for (int i = 0; i < n; i++)
        if (std::fabs(x[i] - xOld[i]) > 1e-12 * std::max(1.0, std::fabs(x[i]))) done = false;
      if (done) break;
    }
    accept(0, ModeOP);
    return E_OK;
  }

  // roundoff, analysis/op.go:275-282
  double roundoff() {
    m.absProduct(x.data(), absBuf.data());
    double sum = 0;
    for (int i = 0; i < n; i++) sum += std::fabs(x[i]) * (absBuf[i] + std::fabs(rhs[i]));
    return 64 * 0x1p-52 * sum;
  }

  // result + BuildLedger + RelativeResidual (analysis/op.go:361-404,
  // em/energy.go:90-119, output/result.go:100-109)
  double ledgerOP(Accumulator* out) {
    Accumulator acc;
    for (auto& d : devs) {
      OperatingPoint op = d->operating(x.data(), n);
      d->ports(op, acc);
      if (d->reactive()) acc.addStorage(op.storage);
      if (d->modelSupplier()) acc.addModelSupplied(d->modelSupplied(op));
    }
    if (out) *out = acc;
    double res = acc.residual();
    double ro = roundoff();
    if (acc.scale == 0 || std::fabs(res) <= std::max(1e-21, ro)) return 0;
    return std::fabs(res / acc.scale);
  }

  // Sim.result's element entries and em.BuildLedger (analysis/op.go:361-404,
  // em/energy.go:89-119): rep[nel][8] = voltage, current, absorbed power,
  // storage power, stored energy, model-supplied power, power through the
  // modeled exterior port and its channel (-1 none); led[10] = inflow,
  // outflow, dissipation, storage rate, model-supplied, hypothetical (0),
  // residual, scale, round-off floor, relative residual.
  void report(double* rep, double* led) {
    for (size_t k = 0; k < devs.size(); k++) {
      Device* d = devs[k].get();
      OperatingPoint op = d->operating(x.data(), n);
      Accumulator probe;
      d->ports(op, probe);
      double* r = rep + k * 8;
      r[0] = op.voltage;
      r[1] = op.current;
      r[2] = op.power;
      r[3] = op.storage;
      r[4] = d->storedEnergy(x.data(), n);
      r[5] = d->modelSupplier() ? d->modelSupplied(op) : 0.0;
      r[6] = probe.lastCh >= 0 ? probe.lastPw : 0.0;
      r[7] = probe.lastCh;
    }
    Accumulator acc;
    double rel = ledgerOP(&acc);
    led[0] = acc.inflow;
    led[1] = acc.outflow;
    led[2] = acc.dissipation;
    led[3] = acc.storageRate;
    led[4] = acc.modelSupplied;
    led[5] = 0;
    led[6] = acc.residual();
    led[7] = acc.scale;
    led[8] = roundoff();
    led[9] = rel;
  }

  // Sim.OP, analysis/op.go:340-357
  int op(const Options& o, int* iters, double* ledgerRel, const double* guess) {
    lc.mode = ModeOP;
    lc.time = 0;
    if (guess) {  // dcSolve with NodeSet: applyNodeSet, InitJct off
      applyGuess(guess);
      lc.initJct = false;
      int err = newton(o.itl1, o, iters);
      if (err != E_OK && !nonlin.empty() && retryable(err)) {
        int first = err;
        err = first;
        if (!o.noGmin && gminStepping(o, iters) == E_OK) err = E_OK;
        else if (!o.noSource && sourceStepping(o, iters) == E_OK) err = E_OK;
        else if (!o.noPT && pseudoTransient(o, iters) == E_OK) err = E_OK;
      }
      if (err != E_OK) return err;
      warm = true;
    } else {
      int e = dcSolve(o, iters);
      if (e != E_OK) return e;
    }
    accept(0, ModeOP);
    double rr = ledgerOP(nullptr);
    if (ledgerRel) *ledgerRel = rr;
    if (!o.skipEnergy && rr > o.ledgerTol) return E_LEDGER;
    return E_OK;
  }

  double nextBreakpoint(double t) {  // analysis/tran.go:233-243
    double nb = kInf;
    for (int i : breakers) {
      double b = devs[i]->nextBreakpoint(t);
      if (b < nb) nb = b;
    }
    return nb;
  }

  // recorder.record's per-point ledger (analysis/tran.go:291-347)
  struct Recorder {
    double peakScale = 0, worstAbs = 0, worstT = 0;
  };
  int record(Recorder& r, double t, bool imposed) {
    Accumulator acc;
    std::vector<OperatingPoint> ops(devs.size());
    for (size_t i = 0; i < devs.size(); i++) ops[i] = devs[i]->operating(x.data(), n);
    for (size_t i = 0; i < devs.size(); i++) devs[i]->ports(ops[i], acc);
    for (size_t i = 0; i < devs.size(); i++)
      if (devs[i]->modelSupplier()) acc.addModelSupplied(devs[i]->modelSupplied(ops[i]));
    for (int i : reactive) acc.addStorage(ops[i].storage);
    int err = E_OK;
    if (acc.scale > r.peakScale) r.peakScale = acc.scale;
    double res = std::fabs(acc.residual());
    if (!imposed && res > std::max(1e-21, roundoff())) {
      if (res > r.worstAbs) { r.worstAbs = res; r.worstT = t; }
      if (res / r.peakScale > 1e-3) err = E_LEDGER;
    }
    return err;
  }

  struct TranOut {
    int accepted = 0, rejected = 0;
    double minStep = kInf, maxStep = 0, maxRel = 0, tEnd = 0;
    int iters = 0;
    // optional trace of accepted points
    double* times = nullptr;
    double* xs = nullptr;
    double* dev_current = nullptr;  // [cap][ndev] element currents at each stored point
    int ndev = 0;
    int cap = 0, count = 0;
  };

  // Sim.tran, analysis/tran.go:95-231 (+ recorder.finish :351-386)
  int tran(const kh_tran_spec& sp, const Options& oo, TranOut& out) {
    Options o = oo;
    int method = sp.method == 1 ? BackwardEuler : Trapezoidal;
    double chgtol = sp.chgtol > 0 ? sp.chgtol : 1e-14;
    double trtol = sp.trtol > 0 ? sp.trtol : 7;
    double minStep = sp.min_step > 0 ? sp.min_step : 1e-12 * sp.stop;
    int64_t maxSteps = sp.max_steps > 0 ? sp.max_steps : 1000000;
    double maxStep = sp.max_step;
    if (maxStep == 0) maxStep = std::min(sp.step, (sp.stop - sp.start) / 50);
    Recorder rec;
    auto store = [&](double t) {
      if (out.times && out.count < out.cap) {
        out.times[out.count] = t;
        std::copy(x.begin(), x.end(), out.xs + (size_t)out.count * n);
        if (out.dev_current)
          for (int k = 0; k < out.ndev; k++)
            out.dev_current[(size_t)out.count * out.ndev + k] = devs[k]->operating(x.data(), n).current;
        out.count++;
      }
    };
    if (sp.uic) {
      for (auto& d : devs) d->initialCondition();
      lc.mode = ModeUICInit;
      lc.time = 0;
      lc.initJct = true;
      int e = newton(o.itl1, o, &out.iters);
      if (e != E_OK) return e;
      accept(0, ModeUICInit);
      record(rec, 0, true);
      store(0);
    } else {
      lc.mode = ModeTranOP;
      lc.time = 0;
      int e = dcSolve(o, &out.iters);
      if (e != E_OK) return e;
      accept(0, ModeTranOP);
      int le = record(rec, 0, false);
      store(0);
      if (le != E_OK && !o.skipEnergy) return le;
    }
    TruncContext tc;
    tc.relTol = o.relTol; tc.absTol = o.absTol; tc.vnTol = o.vnTol; tc.chgTol = chgtol; tc.trTol = trtol;
    double t = 0;
    double h = std::min(maxStep, std::min(sp.step, sp.stop / 100)) / 10;
    int order = BackwardEuler;
    double eps = 1e-12 * sp.stop;
    std::vector<double> xAcc = x;
    auto reject = [&](double hNew) -> int {
      out.rejected++;
      x = xAcc;
      if (hNew < minStep) return E_TSTEP;
      return E_OK;
    };
    while (t < sp.stop - eps) {
      if (out.accepted >= maxSteps) { out.tEnd = t; return E_MAXSTEPS; }
      double nb = nextBreakpoint(t);
      double tn = t + h;
      bool hitBreak = false;
      if (tn >= nb - eps) { tn = nb; hitBreak = true; }
      if (tn >= sp.stop - eps) { tn = sp.stop; hitBreak = false; }
      h = tn - t;
      lc.mode = ModeTran; lc.time = tn; lc.H = h; lc.method = order;
      int e = newton(o.itl4, o, &out.iters);
      if (e != E_OK) {
        if (!retryable(e) || nonlin.empty()) { out.tEnd = t; return e; }
        h /= 8;
        if (reject(h) != E_OK) { out.tEnd = t; return E_TSTEP; }
        order = BackwardEuler;
        continue;
      }
      tc.time = tn; tc.H = h; tc.method = order;
      double proposal = kInf;
      for (int i : reactive) {
        double p = devs[i]->truncate(x.data(), n, tc);
        if (p < proposal) proposal = p;
      }
      if (proposal < 0.9 * h) {
        h = std::max(proposal, h / 8);
        if (reject(h) != E_OK) { out.tEnd = t; return E_TSTEP; }
        continue;
      }
      accept(tn, ModeTran);
      int le = record(rec, tn, false);
      store(tn);
      if (le != E_OK && !o.skipEnergy) { out.tEnd = tn; return le; }
      xAcc = x;
      out.accepted++;
      out.minStep = std::min(out.minStep, h);
      out.maxStep = std::max(out.maxStep, h);
      t = tn;
      double next = std::min(proposal, 2 * h);
      next = std::min(next, maxStep);
      order = method;
      if (hitBreak) {
        order = BackwardEuler;
        double nb2 = nextBreakpoint(t);
        next = std::min(next, 0.1 * std::min(maxStep, nb2 - t));
      }
      h = std::max(next, minStep);
    }
    out.tEnd = t;
    if (rec.peakScale > 0) {
      out.maxRel = rec.worstAbs / rec.peakScale;
      if (out.maxRel > 1e-9 && !o.skipEnergy) return E_LEDGER;
    }
    return E_OK;
  }

  // DCSweep.values, analysis/dc.go:31-46: start to stop inclusive by step.
  static int sweepValues(double start, double stop, double step, std::vector<double>& out) {
    if (step == 0 || std::isnan(start + stop + step) || std::isinf(start + stop + step)) return E_BADSPEC;
    double cnt = (stop - start) / step;
    if (cnt < 0 || cnt > 1000000) return E_BADSPEC;
    int count = (int)std::floor(cnt + 1e-9) + 1;
    out.resize(count);
    for (int k = 0; k < count; k++) out[k] = start + (double)k * step;
    return E_OK;
  }

  SourceValue* sourceOf(int e) {
    Device* d = devs[e].get();
    if (d->kind == KH_ELEM_V) return &static_cast<VoltageSource*>(d)->src;
    if (d->kind == KH_ELEM_I) return &static_cast<CurrentSource*>(d)->src;
    return nullptr;
  }

  // DC, analysis/dc.go:51-130: sweeps[0] is the inner loop; at every point
  // the swept sources' DC values are overridden and Sim.OP runs on the same
  // Sim, so every point after the first starts warm from the previous one.
  // x of each point goes to xout[point][n], its Newton iterations to
  // iters[point]; *done counts the points that succeeded.
  int dcSweep(int nsw, const int* elem, const double* start, const double* stop, const double* step,
              const Options& o, double* xout, int* iters, int* done, double* maxRel, double* iout = nullptr) {
    if (nsw < 1 || nsw > 2) return E_BADSPEC;
    std::vector<double> vals[2];
    SourceValue* srcs[2] = {nullptr, nullptr};
    size_t total = 1;
    for (int i = 0; i < nsw; i++) {
      int e = sweepValues(start[i], stop[i], step[i], vals[i]);
      if (e != E_OK) return e;
      total *= vals[i].size();
      if (elem[i] < 0 || elem[i] >= (int)devs.size() || !(srcs[i] = sourceOf(elem[i]))) return E_BADSPEC;
    }
    if (total > 1000000) return E_BADSPEC;
    std::vector<double> outer = nsw == 2 ? vals[1] : std::vector<double>{0.0};
    *done = 0;
    *maxRel = 0;
    int point = 0, ret = E_OK;
    for (double vo : outer) {
      if (nsw == 2) { srcs[1]->override_ = vo; srcs[1]->overridden = true; }
      for (double vi : vals[0]) {
        srcs[0]->override_ = vi;
        srcs[0]->overridden = true;
        int it = 0;
        double lr = 0;
        int e = op(o, &it, &lr, nullptr);
        if (iters) iters[point] = it;
        if (e != E_OK) { ret = e; goto out; }
        if (xout) std::copy(x.begin(), x.end(), xout + (size_t)point * n);
        if (iout)  // DCSweepResult.I: the elements' currents after the point's Accept (analysis/dc.go:121-123)
          for (size_t k = 0; k < devs.size(); k++)
            iout[(size_t)point * devs.size() + k] = devs[k]->operating(x.data(), n).current;
        if (lr > *maxRel) *maxRel = lr;
        ++*done;
        ++point;
      }
    }
  out:
    for (int i = 0; i < nsw; i++) srcs[i]->overridden = false;  // ClearDCOverride
    return ret;
  }

  // AC sweep, analysis/ac.go:76-193: OP + polish, then per frequency LoadAC,
  // complex GP solve and complex-Tellegen check.
  // Optionally the element currents iout [nf][nel] and the per-frequency
  // ledgers lout [nf][6]: inflow, outflow, dissipation, reactive, model-
  // supplied, residual (analysis/ac.go:118-173).
  int acSweep(const double* freqs, int nf, const Options& o, cplx* xout, double* maxRel, int* iters,
              cplx* iout = nullptr, double* lout = nullptr) {
    double lr;
    int e = op(o, iters, &lr, nullptr);
    if (e != E_OK) return e;
    e = polish();
    if (e != E_OK) return e;
    ComplexMatrix cm;
    if (cm.init(n, resolve_sparse(backendSel, n)) != OK) return E_BADSIZE;
    std::vector<cplx> crhs(n), cx(n);
    std::vector<double> ab(n);
    ACContext actx;
    actx.M = &cm;
    actx.RHS = &crhs;
    *maxRel = 0;
    for (int k = 0; k < nf; k++) {
      cm.clear();
      std::fill(crhs.begin(), crhs.end(), cplx(0));
      actx.omega = 2 * kPi * freqs[k];
      for (auto& d : devs) d->loadAC(actx);
      int se = cm.solve(crhs.data(), cx.data());
      if (se != OK) return se;
      std::copy(cx.begin(), cx.end(), xout + (size_t)k * n);
      cplx sumS = 0;
      double scaleS = 0;
      Accumulator acc;
      double reactive = 0;
      // the average-power ledger of one element's absorbed complex power s
      auto book = [&](size_t kd, Device* d, cplx s, cplx cur) {
        if (iout) iout[(size_t)k * devs.size() + kd] = cur;
        const int kind = d->kind;
        if (kind == KH_ELEM_C || kind == KH_ELEM_L || kind == KH_ELEM_D || kind == KH_ELEM_M || kind == KH_ELEM_Q)
          reactive += s.imag();
        OperatingPoint op;
        op.power = s.real();
        d->ports(op, acc);
        if (d->modelSupplier()) acc.addModelSupplied(d->modelSupplied(op));
      };
      for (size_t kd = 0; kd < devs.size(); kd++) {
        auto& d = devs[kd];
        cplx v, i;
        if (d->kind == KH_ELEM_M) {
          // ACPower, device/mosfet.go:841-846 via acCurrents :817-832
          Mosfet* q = static_cast<Mosfet*>(d.get());
          auto cat = [&](int idx) { return (idx < 0 || idx >= n) ? cplx(0) : cx[idx]; };
          cplx gg = cat(q->g), bb = cat(q->b), dd = cat(q->dp), ss = cat(q->sp);
          cplx vgs = gg - ss, vds = dd - ss, vbs = bb - ss;
          cplx ich = cplx(q->ac.gch[0], 0) * vgs + cplx(q->ac.gch[1], 0) * vds + cplx(q->ac.gch[2], 0) * vbs;
          cplx ibs = cplx(q->ac.gbs, 0) * vbs;
          cplx ibd = cplx(q->ac.gbd, 0) * (bb - dd);
          cplx jw(0, q->ac.omega);
          cplx iqs[5] = {jw * q->ac.cap[0] * vgs, jw * q->ac.cap[1] * (gg - dd), jw * q->ac.cap[2] * (gg - bb),
                         jw * q->ac.cap[3] * vbs, jw * q->ac.cap[4] * (bb - dd)};
          cplx id = ich - ibd - iqs[1] - iqs[4];
          cplx is = -ich - ibs - iqs[0] - iqs[3];
          cplx ig = iqs[0] + iqs[1] + iqs[2];
          cplx ib = ibs + ibd - iqs[2] + iqs[3] + iqs[4];
          cplx s = cat(q->u[0]) * std::conj(id) + cat(q->g) * std::conj(ig) + cat(q->u[1]) * std::conj(is) +
                   cat(q->b) * std::conj(ib);
          s /= 2.0;
          sumS += s;
          scaleS += absval(s);
          book(kd, d.get(), s, id);
          continue;
        }
        if (d->kind == KH_ELEM_Q) {
          cplx ic;
          cplx s = static_cast<Bjt*>(d.get())->acPower(cx.data(), n, &ic);
          sumS += s;
          scaleS += absval(s);
          book(kd, d.get(), s, ic);
          continue;
        }
        v = d->voltageAC(cx.data());
        switch (d->kind) {
          case KH_ELEM_R: i = v * cplx(static_cast<Resistor*>(d.get())->g, 0); break;
          case KH_ELEM_C: i = v * cplx(0, static_cast<Capacitor*>(d.get())->omega * static_cast<Capacitor*>(d.get())->C); break;
          case KH_ELEM_L: i = cx[static_cast<Inductor*>(d.get())->branch]; break;
          case KH_ELEM_V: i = cx[static_cast<VoltageSource*>(d.get())->branch]; break;
          case KH_ELEM_I: i = static_cast<CurrentSource*>(d.get())->src.ac; break;
          case KH_ELEM_D: {
            Diode* dd = static_cast<Diode*>(d.get());
            cplx vj = 0;
            if (dd->ai >= 0 && dd->ai < n) vj = cx[dd->ai];
            if (dd->u[1] >= 0) vj -= cx[dd->u[1]];
            double id, g0, q, c;
            dd->current(dd->vdAcc, &id, &g0);
            dd->charge(dd->vdAcc, id, g0, &q, &c);
            i = vj * cplx(g0 + dd->gmin, dd->omega * c);
            break;
          }
          case KH_ELEM_E:
          case KH_ELEM_F:
          case KH_ELEM_G:
          case KH_ELEM_H: i = static_cast<Controlled*>(d.get())->currentAC(cx.data()); break;
          default: i = 0;
        }
        cplx s = 0.5 * v * std::conj(i);
        sumS += s;
        scaleS += absval(s);
        book(kd, d.get(), s, i);
      }
      if (lout) {
        double* l = lout + (size_t)k * 6;
        l[0] = acc.inflow;
        l[1] = acc.outflow;
        l[2] = acc.dissipation;
        l[3] = reactive;
        l[4] = acc.modelSupplied;
        l[5] = acc.residual();
      }
      cm.absProduct(cx.data(), ab.data());
      double ro = 0;
      for (int i = 0; i < n; i++) ro += absval(cx[i]) * (ab[i] + absval(crhs[i]));
      ro *= 64 * 0x1p-52;
      double rr = 0;
      if (scaleS > 0 && absval(sumS) > ro) rr = absval(sumS) / scaleS;
      if (rr > *maxRel) *maxRel = rr;
      if (!o.skipEnergy && rr > o.ledgerTol) return E_LEDGER;
    }
    return E_OK;
  }
};

}  // namespace oracle

using namespace oracle;

// ---- C API (ctypes) --------------------------------------------------------
extern "C" {

typedef struct oracle_op_info {
  int32_t n;
  int32_t iters;
  int64_t solves;
  int64_t nnz_a, nnz_l, nnz_u;
  int32_t analyses;
  int32_t err_step;
  double ledger_rel;
  double inflow, outflow, dissipation, storage;
} oracle_op_info;

// Size of the MNA system of a circuit (number of unknowns).
int oracle_size(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, 2, 0, 0);
  return e == E_OK ? s.n : -e;
}

// analysis.OP on one circuit. q and prow (optional, n each) receive the final
// factorization's ordering and pivot rows (sparse backend only).
int oracle_op(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params, const kh_options* ko,
              int32_t backend, const double* guess, double* x_out, int32_t* q_out, int32_t* prow_out,
              double* elem_current, oracle_op_info* info) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, backend, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  int iters = 0;
  double lr = 0;
  e = s.op(o, &iters, &lr, guess);
  if (x_out) std::copy(s.x.begin(), s.x.end(), x_out);
  if (elem_current && (e == E_OK || e == E_LEDGER))
    for (int k = 0; k < nel; k++) elem_current[k] = s.devs[k]->operating(s.x.data(), s.n).current;
  if (info) {
    info->n = s.n;
    info->iters = iters;
    info->solves = s.solves;
    info->ledger_rel = lr;
    info->analyses = s.m.sparse ? s.m.s.analyses : 0;
    info->err_step = s.m.sparse ? s.m.s.errStep : s.m.d.errStep;
    if (s.m.sparse) s.m.s.nnz(&info->nnz_a, &info->nnz_l, &info->nnz_u);
    Accumulator acc;
    if (e == E_OK || e == E_LEDGER) {
      s.ledgerOP(&acc);
      info->inflow = acc.inflow;
      info->outflow = acc.outflow;
      info->dissipation = acc.dissipation;
      info->storage = acc.storageRate;
    }
  }
  if (s.m.sparse) {
    if (q_out) std::copy(s.m.s.q.begin(), s.m.s.q.end(), q_out);
    if (prow_out) std::copy(s.m.s.prow.begin(), s.m.s.prow.end(), prow_out);
  }
  return e;
}

// OP then polish (analysis/op.go:243-271): the operating point exact to
// round-off. iters_out counts every solve, the polish's included.
int oracle_op_polished(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params,
                       const kh_options* ko, const double* guess, double* x_out, int32_t* iters_out) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, ko ? ko->backend : 0, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  int iters = 0;
  double lr;
  e = s.op(o, &iters, &lr, guess);
  if (e == E_OK) e = s.polish(&iters);
  if (x_out) std::copy(s.x.begin(), s.x.end(), x_out);
  if (iters_out) *iters_out = iters;  // Newton iterations and polish solves: every solve of the circuit
  return e;
}

// OP (polished when polish != 0), then output.Result's element entries and
// ledger (Sim::report): rep [nel*8], led [10]; x_out [n]. The entries are
// filled whenever the operating point exists (status OK or ledger).
int oracle_op_report(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params,
                     const kh_options* ko, int32_t polish, double* x_out, double* rep, double* led) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, ko ? ko->backend : 0, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  int iters = 0;
  double lr;
  e = s.op(o, &iters, &lr, nullptr);
  if (e == E_OK && polish) e = s.polish(&iters);
  if (x_out) std::copy(s.x.begin(), s.x.end(), x_out);
  if (e == E_OK || e == E_LEDGER) s.report(rep, led);
  return e;
}

// One DC method alone on a freshly built circuit, as
// analysis/newton_test.go:74-104 (TestSteppingAgrees) runs them: 0 dcSolve
// (plain Newton, then the fallbacks the options allow), 1 gminStepping,
// 2 sourceStepping, 3 pseudoTransient (analysis/newton.go:95-311). The
// circuit starts in ModeOP with InitJct off, as a new Sim does.
// trace (optional, cap doubles): the continuation trace of Sim::trace;
// *trace_len receives its length (which may exceed cap: then it is cut).
int oracle_dc_method(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params,
                     const kh_options* ko, int32_t method, double* x_out, int32_t* iters_out, double* trace,
                     int32_t cap, int32_t* trace_len) {
  Sim s;
  std::vector<double> tr;
  if (trace) s.trace = &tr;
  int e = s.build(num_nodes, el, nel, params, ko ? ko->backend : 0, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  s.lc.mode = ModeOP;
  s.lc.time = 0;
  int iters = 0;
  switch (method) {
    case 0: e = s.dcSolve(o, &iters); break;
    case 1: e = s.gminStepping(o, &iters); break;
    case 2: e = s.sourceStepping(o, &iters); break;
    case 3: e = s.pseudoTransient(o, &iters); break;
    default: return E_BADSPEC;
  }
  if (x_out) std::copy(s.x.begin(), s.x.end(), x_out);
  if (iters_out) *iters_out = iters;
  if (trace) {
    for (size_t k = 0; k < tr.size() && (int32_t)k < cap; k++) trace[k] = tr[k];
    if (trace_len) *trace_len = (int32_t)tr.size();
  }
  return e;
}

// Sets a planted mistake (tests only): which 1 gmin factor cap, 2 source-step
// growth, 3 pseudo-transient growth; value 0 restores the reference's constant.
void oracle_plant_stepping(int32_t which, double value) {
  if (which == 1) g_plant_gmin_cap = value;
  if (which == 2) g_plant_src_growth = value;
  if (which == 3) g_plant_pt_growth = value;
}

// lteStep (device/lte.go:50-85) on an explicit history: candidate point
// (time, qn), accepted points ts[0..2], qs[0..2] (most recent first), hist_n
// valid; method 1 backward Euler, 2 trapezoidal.
double oracle_lte_step(int32_t method, double time, double H, const double* ts, const double* qs, int32_t hist_n,
                       double qn, double in, double in_prev, double abs_tol, double rel_tol, double chg_tol,
                       double tr_tol) {
  History h;
  for (int k = 0; k < 3; k++) { h.t[k] = ts[k]; h.q[k] = qs[k]; }
  h.n = hist_n;
  TruncContext tc;
  tc.time = time;
  tc.H = H;
  tc.method = method;
  tc.relTol = rel_tol;
  tc.absTol = abs_tol;
  tc.chgTol = chg_tol;
  tc.trTol = tr_tol;
  return lteStep(h, qn, in, in_prev, abs_tol, tc);
}

typedef struct oracle_tran_info {
  int32_t n, accepted, rejected, iters, count;
  double min_step, max_step, max_rel, t_end;
} oracle_tran_info;

int oracle_tran(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params, const kh_options* ko,
                const kh_tran_spec* spec, int32_t backend, double* x_out, double* times, double* xs, double* cur,
                int32_t cap, oracle_tran_info* info) {
  Sim s;
  if (!(spec->step > 0) || !(spec->stop > 0) || spec->start < 0 || spec->start >= spec->stop || spec->max_step < 0)
    return E_BADSPEC;
  int e = s.build(num_nodes, el, nel, params, backend, spec->step, spec->stop);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  Sim::TranOut out;
  out.times = times;
  out.xs = xs;
  out.cap = cap;
  out.dev_current = cur;
  out.ndev = nel;
  e = s.tran(*spec, o, out);
  if (x_out) std::copy(s.x.begin(), s.x.end(), x_out);
  if (info) {
    info->n = s.n;
    info->accepted = out.accepted;
    info->rejected = out.rejected;
    info->iters = out.iters;
    info->count = out.count;
    info->min_step = out.minStep;
    info->max_step = out.maxStep;
    info->max_rel = out.maxRel;
    info->t_end = out.tEnd;
  }
  return e;
}

// DC sweep of one or two sources (element indices src[0] inner, src[1]
// outer); xout [points][n], iters [points] (both sized by the caller for every
// point of the sweep), *done = points solved.
int oracle_dc(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params, const kh_options* ko,
              int32_t nsweeps, const int32_t* src, const double* start, const double* stop, const double* step,
              double* xout, int32_t* iters, int32_t* done, double* max_rel, double* iout) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, ko ? ko->backend : 0, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  int el2[2] = {src[0], nsweeps > 1 ? src[1] : -1};
  int d = 0;
  e = s.dcSweep(nsweeps, el2, start, stop, step, o, xout, iters, &d, max_rel, iout);
  *done = d;
  return e;
}

// AC sweep; xout is [nf][n] complex128 (interleaved re, im).
int oracle_ac(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params, const kh_options* ko,
              const double* freqs, int32_t nf, double* xout, double* max_rel) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, ko ? ko->backend : 0, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  int iters = 0;
  return s.acSweep(freqs, nf, o, reinterpret_cast<cplx*>(xout), max_rel, &iters);
}

// oracle_ac with ACResult's element currents iout [nf][nel] (re, im) and the
// per-frequency ledgers lout [nf][6].
int oracle_ac_ex(int32_t num_nodes, const kh_element* el, int32_t nel, const double* params, const kh_options* ko,
                 const double* freqs, int32_t nf, double* xout, double* max_rel, double* iout, double* lout) {
  Sim s;
  int e = s.build(num_nodes, el, nel, params, ko ? ko->backend : 0, 0, 0);
  if (e != E_OK) return e;
  Options o;
  o.from(ko);
  int iters = 0;
  return s.acSweep(freqs, nf, o, reinterpret_cast<cplx*>(xout), max_rel, &iters, reinterpret_cast<cplx*>(iout), lout);
}

// ---- raw sparse-LU access (solver.Sparse seam) for golden-vector tests -------
// ---- JacobianTearing (kron/jacobian.go:66-712) --------------------------------
// A Tearing of the given cut pairs (node unknowns, -1 ground; [2*ncuts]), its
// torn matrices, and the switch that makes the analyses' matrices torn ones
// (Options.Solver). Single-threaded: the counters are plain integers.
void* oracle_tear_new(const int32_t* cuts, int32_t ncuts, double reuse_tol) {
  auto* t = new Tearing();
  for (int k = 0; k < ncuts; k++) t->cuts.push_back({cuts[2 * k], cuts[2 * k + 1]});
  t->reuseTol = reuse_tol;
  return t;
}
void oracle_tear_free(void* t) {
  if (g_tear == t) g_tear = nullptr;
  delete static_cast<Tearing*>(t);
}
void oracle_tear_use(void* t) { g_tear = static_cast<Tearing*>(t); }
// solves, factors, reuses, stale reuses, fallbacks, refinements, max pieces
void oracle_tear_stats(void* h, int64_t* out) {
  const Tearing* t = static_cast<const Tearing*>(h);
  out[0] = t->solves; out[1] = t->factors; out[2] = t->reuses; out[3] = t->stale;
  out[4] = t->fallbacks; out[5] = t->refines; out[6] = t->maxPieces;
}
void* oracle_tmat_new(void* tear, int32_t n, int32_t cplx_) {
  if (cplx_) {
    auto* m = new TornMat<cplx>();
    m->init(n, static_cast<Tearing*>(tear));
    return m;
  }
  auto* m = new TornMat<double>();
  m->init(n, static_cast<Tearing*>(tear));
  return m;
}
void oracle_tmat_free(void* h, int32_t cplx_) {
  if (cplx_) delete static_cast<TornMat<cplx>*>(h);
  else delete static_cast<TornMat<double>*>(h);
}
void oracle_tmat_clear(void* h, int32_t cplx_) {
  if (cplx_) static_cast<TornMat<cplx>*>(h)->clear();
  else static_cast<TornMat<double>*>(h)->clear();
}
// v holds (re, im) pairs when complex
void oracle_tmat_add_batch(void* h, int32_t cplx_, const int32_t* i, const int32_t* j, const double* v, int64_t count) {
  for (int64_t k = 0; k < count; k++) {
    if (cplx_) static_cast<TornMat<cplx>*>(h)->add(i[k], j[k], cplx(v[2 * k], v[2 * k + 1]));
    else static_cast<TornMat<double>*>(h)->add(i[k], j[k], v[k]);
  }
}
int oracle_tmat_solve(void* h, int32_t cplx_, const double* b, double* x) {
  if (cplx_)
    return static_cast<TornMat<cplx>*>(h)->solve(reinterpret_cast<const cplx*>(b), reinterpret_cast<cplx*>(x));
  return static_cast<TornMat<double>*>(h)->solve(b, x);
}
void oracle_tmat_absproduct(void* h, int32_t cplx_, const double* x, double* y) {
  if (cplx_) static_cast<TornMat<cplx>*>(h)->absProduct(reinterpret_cast<const cplx*>(x), y);
  else static_cast<TornMat<double>*>(h)->absProduct(x, y);
}
// Recording of torn solves (TearRecord): on / off (off drops the records),
// their count, and one record's sizes and contents.
static std::vector<TearRecord> g_records;
void oracle_tear_record(int32_t on) {
  g_records.clear();
  g_tear_rec = on ? &g_records : nullptr;
}
int64_t oracle_tear_rec_count() { return (int64_t)g_records.size(); }
void oracle_tear_rec_info(int64_t k, int32_t* mat, int32_t* cplx_, int32_t* n, int64_t* nslots, int32_t* status) {
  const TearRecord& r = g_records[k];
  *mat = r.mat; *cplx_ = r.cplx; *n = r.n; *nslots = (int64_t)r.ri.size(); *status = r.status;
}
void oracle_tear_rec_get(int64_t k, int32_t* ri, int32_t* ci, double* val, double* b, double* x) {
  const TearRecord& r = g_records[k];
  std::copy(r.ri.begin(), r.ri.end(), ri);
  std::copy(r.ci.begin(), r.ci.end(), ci);
  std::copy(r.val.begin(), r.val.end(), val);
  std::copy(r.b.begin(), r.b.end(), b);
  std::copy(r.x.begin(), r.x.end(), x);
}
// Pieces of the held plan (-1: none), and the piece of each unknown [n].
int oracle_tmat_pieces(void* h, int32_t cplx_, int32_t* piece_of) {
  auto get = [&](auto* m) -> int {
    if (!m->plan) return -1;
    if (piece_of) for (int i = 0; i < m->n; i++) piece_of[i] = m->plan->pieceOf[i];
    return (int)m->plan->pieces.size();
  };
  if (cplx_) return get(static_cast<TornMat<cplx>*>(h));
  return get(static_cast<TornMat<double>*>(h));
}

void* oracle_sparse_new(int32_t n, int32_t cplx_) {
  if (cplx_) {
    auto* s = new SparseLU<cplx>();
    if (s->init(n) != OK) { delete s; return nullptr; }
    return s;
  }
  auto* s = new SparseLU<double>();
  if (s->init(n) != OK) { delete s; return nullptr; }
  return s;
}
void oracle_sparse_free(void* h, int32_t cplx_) {
  if (cplx_) delete static_cast<SparseLU<cplx>*>(h);
  else delete static_cast<SparseLU<double>*>(h);
}
void oracle_sparse_clear(void* h, int32_t cplx_) {
  if (cplx_) static_cast<SparseLU<cplx>*>(h)->clear();
  else static_cast<SparseLU<double>*>(h)->clear();
}
void oracle_sparse_add(void* h, int32_t i, int32_t j, double re, double im, int32_t cplx_) {
  if (cplx_) static_cast<SparseLU<cplx>*>(h)->add(i, j, cplx(re, im));
  else static_cast<SparseLU<double>*>(h)->add(i, j, re);
}
int oracle_sparse_solve(void* h, const double* b, double* x, int32_t cplx_) {
  if (cplx_) {
    auto* s = static_cast<SparseLU<cplx>*>(h);
    return s->solve(reinterpret_cast<const cplx*>(b), reinterpret_cast<cplx*>(x));
  }
  return static_cast<SparseLU<double>*>(h)->solve(b, x);
}
void oracle_sparse_info(void* h, int32_t cplx_, int64_t* a, int64_t* l, int64_t* u, int32_t* q, int32_t* prow,
                        int32_t* err_step) {
  if (cplx_) {
    auto* s = static_cast<SparseLU<cplx>*>(h);
    s->nnz(a, l, u);
    if (q) std::copy(s->q.begin(), s->q.end(), q);
    if (prow) std::copy(s->prow.begin(), s->prow.end(), prow);
    if (err_step) *err_step = s->errStep;
    return;
  }
  auto* s = static_cast<SparseLU<double>*>(h);
  s->nnz(a, l, u);
  if (q) std::copy(s->q.begin(), s->q.end(), q);
  if (prow) std::copy(s->prow.begin(), s->prow.end(), prow);
  if (err_step) *err_step = s->errStep;
}
void oracle_sparse_absproduct(void* h, const double* x, double* y) {
  static_cast<SparseLU<double>*>(h)->absProduct(x, y);
}
// CSparse.AbsProduct (solver/sparse.go:519-527); x is n (re, im) pairs.
void oracle_csparse_absproduct(void* h, const double* x, double* y) {
  static_cast<SparseLU<cplx>*>(h)->absProduct(reinterpret_cast<const cplx*>(x), y);
}
void oracle_min_degree(int32_t n, const int32_t* ri, const int32_t* ci, int64_t nnz, int32_t* order) {
  std::vector<int32_t> r(ri, ri + nnz), c(ci, ci + nnz);
  std::vector<int> o = minimumDegree(n, r, c);
  for (int k = 0; k < n; k++) order[k] = o[k];
}

// Dense backend (solver.Dense) for cross-checks.
int oracle_dense_solve(int32_t n, const int32_t* ri, const int32_t* ci, const double* v, int64_t nnz, const double* b,
                       double* x) {
  Dense d;
  if (d.init(n) != OK) return BADSIZE;
  for (int64_t k = 0; k < nnz; k++) d.add(ri[k], ci[k], v[k]);
  return d.solve(b, x);
}

}

// ---- batch runners: one worker thread per core, one task per circuit --------
// (the evaluation pool of evo/engine.go:519-544; results stored by index).
static void run_pool_impl(int B, int threads, const std::function<void(int)>& f) {
  std::atomic<int> next(0);
  std::vector<std::thread> ts;
  if (threads < 1) threads = 1;
  for (int t = 0; t < threads; t++)
    ts.emplace_back([&]() {
      for (;;) {
        int c = next.fetch_add(1);
        if (c >= B) break;
        f(c);
      }
    });
  for (auto& t : ts) t.join();
}

extern "C" {
// OP for B circuits of one topology; params SoA [nparams][B]; x_out SoA [n][B].
int oracle_batch_op(int32_t num_nodes, const kh_element* el, int32_t nel, int32_t nparams, int32_t B,
                    const double* params, const kh_options* ko, int32_t threads, int32_t polish, double* x_out,
                    int32_t* status, int32_t* iters) {
  run_pool_impl(B, threads, [&](int c) {
    std::vector<double> p(nparams);
    for (int k = 0; k < nparams; k++) p[k] = params[(size_t)k * B + c];
    Sim s;
    int e = s.build(num_nodes, el, nel, p.data(), ko ? ko->backend : 0, 0, 0);
    int it = 0;
    if (e == E_OK) {
      Options o;
      o.from(ko);
      double lr;
      e = s.op(o, &it, &lr, nullptr);
      if (e == E_OK && polish) e = s.polish();
    }
    if (status) status[c] = e;
    if (iters) iters[c] = it;
    if (x_out)
      for (int i = 0; i < s.n; i++) x_out[(size_t)i * B + c] = s.x[i];
  });
  return 0;
}

int oracle_batch_tran(int32_t num_nodes, const kh_element* el, int32_t nel, int32_t nparams, int32_t B,
                      const double* params, const kh_options* ko, const kh_tran_spec* spec, const double* lane_spec,
                      int32_t threads, double* x_out, int32_t* status, int32_t* accepted, int32_t* rejected,
                      double* ledger) {
  run_pool_impl(B, threads, [&](int c) {
    std::vector<double> p(nparams);
    for (int k = 0; k < nparams; k++) p[k] = params[(size_t)k * B + c];
    Sim s;
    kh_tran_spec sp = *spec;
    if (lane_spec) {  // per-circuit step, stop, max step, min step
      sp.step = lane_spec[c];
      sp.stop = lane_spec[(size_t)B + c];
      sp.max_step = lane_spec[(size_t)2 * B + c];
      sp.min_step = lane_spec[(size_t)3 * B + c];
    }
    int e = s.build(num_nodes, el, nel, p.data(), ko ? ko->backend : 0, sp.step, sp.stop);
    Sim::TranOut out;
    if (e == E_OK) {
      Options o;
      o.from(ko);
      e = s.tran(sp, o, out);
    }
    if (status) status[c] = e;
    if (accepted) accepted[c] = out.accepted;
    if (rejected) rejected[c] = out.rejected;
    if (ledger) ledger[c] = out.maxRel;
    if (x_out)
      for (int i = 0; i < s.n; i++) x_out[(size_t)i * B + c] = s.x[i];
  });
  return 0;
}
}
