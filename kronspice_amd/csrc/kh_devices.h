// kh_devices.h — per-lane device models: linearize, stamp, accept, LTE and
// ledger terms for the elements on the solve path.
//
// Each function restates one reference device method (cited inline) for one
// circuit (lane c). Parameters, device state, matrix values and vectors are
// SoA arrays with stride B. Stamps go through the element's slot table:
// add(ti, tj, v) accumulates v into the value slot of the unknown pair
// (term[ti], term[tj]); pairs with a ground index have slot -1 and are
// dropped, as solver.Sparse.Add drops negative indices
// (solver/sparse.go:114-117). Accumulation happens in the reference's Add
// order, so every slot value is bit-identical to the reference's.
#pragma once
#include <stdint.h>
#include <math.h>
#include "kh_lu.h"
#include "kh_mos23.h"

namespace kh {

// Element table layout (int32 words per element): up to KH_NT terminals
// (external nodes, branches and internal nodes; the BJT has 7).
constexpr int KH_NT = 8;
enum {
  ET_KIND = 0,
  ET_TERM = 1,                          // KH_NT unknown indices (-1 = ground / unused)
  ET_POFF = ET_TERM + KH_NT,
  ET_SOFF = ET_POFF + 1,
  ET_SLOT = ET_SOFF + 1,                // KH_NT x KH_NT slot table
  ET_W = ET_SLOT + KH_NT * KH_NT
};

enum { K_R = 1, K_C = 2, K_L = 3, K_V = 4, K_I = 5, K_D = 6, K_M = 7, K_Q = 8 };
// K_M23: not an element kind, the bit of the MOSFET level 2 and 3 currents.
enum { K_M23 = 9 };
// Linear controlled sources: E (VCVS), F (CCCS), G (VCCS), H (CCVS).
enum { K_E = 10, K_F = 11, K_G = 12, K_H = 13 };
KH_HD inline bool is_controlled(int kind) { return kind >= K_E && kind <= K_H; }
// Mutual inductance between two inductors (SPICE K).
enum { K_K = 14 };

// Element kinds a translation unit's circuit code handles (bit k: kind k).
// The persistent kernel is compiled once per set, so a batch runs code for
// the devices it has and nothing else (smaller code, fewer registers).
#define KH_KINDS_ALL 32766
#ifndef KH_KINDS
#define KH_KINDS KH_KINDS_ALL
#endif
#define KH_HAS(k) ((((unsigned)(KH_KINDS)) >> (k)) & 1u)
#define KH_CAT2(a, b) a##b
#define KH_CAT(a, b) KH_CAT2(a, b)
enum { MODE_OP = 0, MODE_TRANOP = 1, MODE_UICINIT = 2, MODE_TRAN = 3 };
enum { BE = 1, TRAP = 2 };

// State sizes (doubles per element).
enum { NS_R = 0, NS_C = 13, NS_L = 9, NS_V = 1, NS_I = 1, NS_D = 19, NS_M = 96, NS_Q = 66, NS_CTL = 0, NS_K = 2 };

// history (device/lte.go:23-37) lives in 7 state words: t0 t1 t2 q0 q1 q2 n.
constexpr double kInf = __builtin_huge_val();
constexpr double kPi = 3.14159265358979323846264338327950288419716939937510582097494459;
constexpr double kE = 2.71828182845904523536028747135266249775724709369995957496696763;

// One circuit's device view: its own parameter, state, value and vector
// arrays, unit stride.
struct Lane {
  int n;
  const int32_t* et;     // element table
  const double* par;     // [P]
  double* st;            // [S]
  double* val;           // [nslots]
  uint8_t* present;      // [nslots]
  double* rhs;           // [n]
  const double* X;       // [n] linearization point (xOld)
  bool grew;
  // Record mode (rslot non-null): every add and addRHS is written to the
  // element's record, in call order, as (target, value) with target = slot,
  // or nslots + row for the right-hand side, instead of being accumulated;
  // the wavefront form sums the records of all elements per target in element
  // order (kh_wave.h). rn counts the calls, rcap bounds the record.
  int32_t* rslot;
  double* rval;
  int32_t rn, rcap, nslots;
  // LoadContext (device/device.go:127-148)
  int mode, method;
  double time, H, gmin, srcFact;
  bool initJct, srcStepping;

  KH_HD double xat(const double* v, int i) const { return (i < 0 || i >= n) ? 0.0 : v[i]; }
  KH_HD double P(const int32_t* e, int k) const { return par[e[ET_POFF] + k]; }
  KH_HD double& S(const int32_t* e, int k) const { return st[e[ET_SOFF] + k]; }
  KH_HD int T(const int32_t* e, int k) const { return e[ET_TERM + k]; }
  KH_HD void add(const int32_t* e, int ti, int tj, double v) {
    int s = e[ET_SLOT + ti * KH_NT + tj];
    if (s < 0) return;
    if (rslot) {
      if (rn < rcap) { rslot[rn] = s; rval[rn] = v; }
      rn++;
      return;
    }
    int64_t ix = s;
    val[ix] = val[ix] + v;
    if (!present[ix]) {
      present[ix] = 1;
      grew = true;
    }
  }
  KH_HD void addRHS(int i, double v) {
    if (i < 0 || i >= n) return;
    if (rslot) {
      if (rn < rcap) { rslot[rn] = nslots + i; rval[rn] = v; }
      rn++;
      return;
    }
    rhs[i] += v;
  }
  KH_HD double volt(const int32_t* e, const double* x) const {  // Base.voltage, device/device.go:310-319
    double v = 0;
    int a = T(e, 0), b = T(e, 1);
    if (a >= 0) v += x[a];
    if (b >= 0) v -= x[b];
    return v;
  }
};

// The most add and addRHS calls one Load of an element of this kind makes
// (load_R ... load_Q below: R 4; C 4 + 2; L 5 + 1; V 4 + 1; I 2; D 8 + 2;
// M: channel 12 + 2, two bulk junctions 2 x (4 + 2), five capacitances
// 5 x (4 + 2), two series resistances 2 x 4; Q: the transport branch 8 + 2,
// three branches 3 x (4 + 2), the charge branches 8 + 2 and 3 x (4 + 2), three
// series resistances 3 x 4): the size of its record in the wavefront form.
KH_HD inline int record_capacity(int kind) {
  switch (kind) {
    case K_R: return 4;
    case K_C: return 6;
    case K_L: return 6;
    case K_V: return 5;
    case K_I: return 2;
    case K_D: return 10;
    case K_M: return 64;
    case K_Q: return 68;
    case K_E: return 6;
    case K_F: return 2;
    case K_G: return 4;
    case K_H: return 5;
    case K_K: return 4;
  }
  return 0;
}

struct OpPt {  // device.OperatingPoint
  double voltage, current, power, storage;
};

// ---- history / lteStep (device/lte.go:29-85) ------------------------------
KH_HD inline void hist_push(const Lane& L, const int32_t* e, int h, double t, double q) {
  L.S(e, h + 2) = L.S(e, h + 1); L.S(e, h + 1) = L.S(e, h + 0); L.S(e, h + 0) = t;
  L.S(e, h + 5) = L.S(e, h + 4); L.S(e, h + 4) = L.S(e, h + 3); L.S(e, h + 3) = q;
  if (L.S(e, h + 6) < 3) L.S(e, h + 6) += 1;
}
KH_HD inline void hist_reset(const Lane& L, const int32_t* e, int h) { L.S(e, h + 6) = 0; }

struct TruncCtx {
  double time, H;
  int method;
  double relTol, absTol, vnTol, chgTol, trTol;
};

KH_HD inline double lte_step(const Lane& L, const int32_t* e, int h, double qn, double in, double inPrev,
                             double absTol, const TruncCtx& tc) {
  int k = tc.method;
  int hn = (int)L.S(e, h + 6);
  if (k < 1 || k > 2 || hn < k + 1) return kInf;
  double ts[4] = {tc.time, L.S(e, h + 0), L.S(e, h + 1), L.S(e, h + 2)};
  double qs[4] = {qn, L.S(e, h + 3), L.S(e, h + 4), L.S(e, h + 5)};
  double dd[4] = {0, 0, 0, 0};
  for (int i = 0; i < k + 2; i++) dd[i] = qs[i];
  for (int order = 1; order <= k + 1; order++) {
    for (int i = 0; i + order <= k + 1; i++) {
      double den = ts[i] - ts[i + order];
      if (den == 0) return kInf;
      dd[i] = (dd[i] - dd[i + 1]) / den;
    }
  }
  double deriv = dd[0];
  double fact = 2.0, cc = 0.5;
  if (k == 2) { fact = 6; cc = 1.0 / 12; }
  double hh = tc.H;
  // math.Pow(h, k) and math.Pow(r, 1/k) for k = 1, 2: Go's Pow returns x for
  // y = 1, Sqrt(x) for y = 0.5 and squares the mantissa exactly once for y = 2
  // (math/pow.go), so h, h*h, r and sqrt(r) are its results bit for bit.
  double errI = cc * (k == 1 ? hh : hh * hh) * fabs(deriv * fact);
  if (errI == 0 || isnan(errI)) return kInf;
  double qmax = fmax(fabs(qn), fabs(L.S(e, h + 3)));
  double allowed = tc.trTol * fmax(tc.relTol * fmax(fabs(in), fabs(inPrev)) + absTol,
                                   tc.relTol * fmax(qmax, tc.chgTol) / hh);
  return hh * (k == 1 ? allowed / errI : sqrt(allowed / errI));
}

// ---- sources (device/source.go:43-66, device/waveform.go:69-239) -----------
KH_HD inline double wave_value(const Lane& L, const int32_t* e, double t) {
  int kind = (int)L.P(e, 2);
  double a0 = L.P(e, 3), a1 = L.P(e, 4), a2 = L.P(e, 5), a3 = L.P(e, 6), a4 = L.P(e, 7), a5 = L.P(e, 8), a6 = L.P(e, 9);
  if (kind == 1) {
    double v1 = a0, v2 = a1, td = a2, tr = a3, tf = a4, pw = a5, per = a6;
    if (t < td) return v1;
    double tt = t - td;
    if (per > 0) tt = fmod(tt, per);
    if (tt < tr) return v1 + (v2 - v1) * tt / tr;
    if (tt < tr + pw) return v2;
    if (tt < tr + pw + tf) return v2 + (v1 - v2) * (tt - tr - pw) / tf;
    return v1;
  }
  if (kind == 2) {
    double vo = a0, va = a1, f = a2, td = a3, theta = a4, ph = a5 * kPi / 180;
    if (t < td) return vo + va * sin(ph);
    double tt = t - td;
    return vo + va * exp(-tt * theta) * sin(2 * kPi * f * tt + ph);
  }
  if (kind == 3) {
    double v1 = a0, v2 = a1, td1 = a2, tau1 = a3, td2 = a4, tau2 = a5;
    if (t < td1) return v1;
    if (t < td2) return v1 + (v2 - v1) * (1 - exp(-(t - td1) / tau1));
    return v1 + (v2 - v1) * (1 - exp(-(t - td1) / tau1)) + (v1 - v2) * (1 - exp(-(t - td2) / tau2));
  }
  return 0;
}

KH_HD inline double wave_next_breakpoint(const Lane& L, const int32_t* e, double t) {
  int kind = (int)L.P(e, 2);
  const double rel = 1e-12;
  double lim = t + rel * fmax(fabs(t), 1e-15);
  if (kind == 1) {
    double td = L.P(e, 5), tr = L.P(e, 6), tf = L.P(e, 7), pw = L.P(e, 8), per = L.P(e, 9);
    double edges[4] = {0, tr, tr + pw, tr + pw + tf};
    if (per <= 0) {
      for (int i = 0; i < 4; i++)
        if (td + edges[i] > lim) return td + edges[i];
      return kInf;
    }
    double k = floor((t - td) / per);
    if (k < 0) k = 0;
    for (double kk = k; kk <= k + 1; kk++) {
      double base = td + kk * per;
      for (int i = 0; i < 4; i++)
        if (base + edges[i] > lim) return base + edges[i];
    }
    return td + (k + 2) * per;
  }
  if (kind == 2) {
    double td = L.P(e, 6);
    if (td > lim) return td;
  }
  if (kind == 3) {
    double td1 = L.P(e, 5), td2 = L.P(e, 7);
    if (td1 > lim) return td1;
    if (td2 > lim) return td2;
  }
  return kInf;
}

KH_HD inline double source_at(const Lane& L, const int32_t* e) {  // sourceValue.at / raw
  double v;
  int kind = (int)L.P(e, 2);
  if (kind == 0) v = L.P(e, 0);
  else if (L.mode == MODE_OP && L.P(e, 1) != 0) v = L.P(e, 0);
  else v = wave_value(L, e, L.mode == MODE_TRAN ? L.time : 0.0);
  if (L.srcStepping && L.mode != MODE_TRAN) v *= L.srcFact;
  return v;
}

// ---- R (device/passive.go:65-104) -----------------------------------------
KH_HD inline void load_R(Lane& L, const int32_t* e) {
  double g = L.P(e, 1);
  L.add(e, 0, 0, g);
  L.add(e, 1, 1, g);
  L.add(e, 0, 1, -g);
  L.add(e, 1, 0, -g);
}

// ---- E, F, G, H (device/controlled.go:76-161) --------------------------------
// terminals: 0 n+, 1 n-, 2 the element's own branch (E, H), 3 nc+ (E, G) or
// the controlling element's branch (F, H), 4 nc- (E, G); parameter 0 the gain.
// Stamps in the reference's order; no state, no right-hand side.
KH_HD inline void load_ctl(Lane& L, const int32_t* e) {
  const double g = L.P(e, 0);
  const int kind = e[ET_KIND];
  if (kind == K_E || kind == K_H) {
    L.add(e, 0, 2, 1);
    L.add(e, 1, 2, -1);
    L.add(e, 2, 0, 1);
    L.add(e, 2, 1, -1);
    if (kind == K_E) {
      L.add(e, 2, 3, -g);
      L.add(e, 2, 4, g);
    } else {
      L.add(e, 2, 3, -g);
    }
    return;
  }
  if (kind == K_G) {
    L.add(e, 0, 3, g);
    L.add(e, 0, 4, -g);
    L.add(e, 1, 3, -g);
    L.add(e, 1, 4, g);
  } else {
    L.add(e, 0, 3, g);
    L.add(e, 1, 3, -g);
  }
}
// Controlled.Operating: the output port's voltage and current (n+ -> n-
// through the element), absorbed power v·i.
KH_HD inline OpPt op_ctl(const Lane& L, const int32_t* e, const double* x) {
  OpPt o;
  const int kind = e[ET_KIND];
  o.voltage = L.volt(e, x);
  if (kind == K_E || kind == K_H) {
    o.current = L.xat(x, L.T(e, 2));
  } else {
    const double ctl = kind == K_G ? L.xat(x, L.T(e, 3)) - L.xat(x, L.T(e, 4)) : L.xat(x, L.T(e, 3));
    o.current = L.P(e, 0) * ctl;
  }
  o.power = o.voltage * o.current;
  o.storage = 0;
  return o;
}

// ---- K (device/mutual.go:74-157) --------------------------------------------
// terminals: 0 the first inductor's branch, 1 the second's; parameter 0 M;
// state: 0 i1Prev, 1 i2Prev. Nothing at DC (the inductors are shorts); in a
// transient step each inductor's integration formula applied to the flux the
// other's current links: -f off the diagonal and -f·i_other,prev on the
// right-hand side, f = M/h (backward Euler) or 2M/h (trapezoidal).
KH_HD inline void load_K(Lane& L, const int32_t* e) {
  if (L.mode != MODE_TRAN) return;
  const double m = L.P(e, 0);
  double f = m / L.H;
  if (L.method == TRAP) f = 2 * m / L.H;
  L.add(e, 0, 1, -f);
  L.add(e, 1, 0, -f);
  L.addRHS(L.T(e, 0), -f * L.S(e, 1));
  L.addRHS(L.T(e, 1), -f * L.S(e, 0));
}
// Mutual.Operating: no voltage, current or power of its own.
KH_HD inline OpPt op_K() {
  OpPt o = {0, 0, 0, 0};
  return o;
}
// Mutual.Accept: the two currents, in every mode.
KH_HD inline void accept_K(const Lane& L, const int32_t* e, const double* x) {
  L.S(e, 0) = x[L.T(e, 0)];
  L.S(e, 1) = x[L.T(e, 1)];
}

// ---- C (device/reactive.go:60-181) ------------------------------------------
// state: 0 geq, 1 ihist, 2 tran, 3 vPrev, 4 iPrev, 5 omega, 6..12 history
KH_HD inline void load_C(Lane& L, const int32_t* e) {
  if (L.mode == MODE_OP || L.mode == MODE_TRANOP) {
    L.S(e, 2) = 0; L.S(e, 0) = 0; L.S(e, 1) = 0;
    return;
  }
  int a = L.T(e, 0), b = L.T(e, 1);
  if (L.mode == MODE_UICINIT) {
    L.S(e, 2) = 0;
    double g = 1e9;
    L.add(e, 0, 0, g); L.add(e, 1, 1, g); L.add(e, 0, 1, -g); L.add(e, 1, 0, -g);
    L.addRHS(a, g * L.S(e, 3));
    L.addRHS(b, -g * L.S(e, 3));
    L.S(e, 0) = 0; L.S(e, 1) = 0;
    return;
  }
  L.S(e, 2) = 1;
  double Cv = L.P(e, 0);
  double geq, ihist;
  if (L.method == TRAP) {
    geq = 2 * Cv / L.H;
    ihist = -(geq * L.S(e, 3) + L.S(e, 4));
  } else {
    geq = Cv / L.H;
    ihist = -geq * L.S(e, 3);
  }
  L.S(e, 0) = geq; L.S(e, 1) = ihist;
  L.add(e, 0, 0, geq); L.add(e, 1, 1, geq); L.add(e, 0, 1, -geq); L.add(e, 1, 0, -geq);
  L.addRHS(a, -ihist);
  L.addRHS(b, ihist);
}
KH_HD inline OpPt op_C(const Lane& L, const int32_t* e, const double* x) {
  OpPt o;
  o.voltage = L.volt(e, x);
  double i = 0;
  if (L.S(e, 2) != 0) i = L.S(e, 0) * o.voltage + L.S(e, 1);
  o.current = i;
  o.power = o.voltage * i;
  o.storage = o.voltage * i;
  return o;
}
KH_HD inline void accept_C(const Lane& L, const int32_t* e, const double* x, double t, int mode) {
  if (mode == MODE_UICINIT) {
    L.S(e, 4) = 0;
    hist_reset(L, e, 6);
  } else if (mode == MODE_OP || mode == MODE_TRANOP) {
    L.S(e, 3) = L.volt(e, x);
    L.S(e, 4) = 0;
    hist_reset(L, e, 6);
  } else {
    OpPt o = op_C(L, e, x);
    L.S(e, 3) = o.voltage;
    L.S(e, 4) = o.current;
  }
  hist_push(L, e, 6, t, L.P(e, 0) * L.S(e, 3));
}
KH_HD inline double trunc_C(const Lane& L, const int32_t* e, const double* x, const TruncCtx& tc) {
  OpPt o = op_C(L, e, x);
  return lte_step(L, e, 6, L.P(e, 0) * o.voltage, o.current, L.S(e, 4), tc.absTol, tc);
}

// ---- L (device/reactive.go:230-349); terms: a, b, branch k ------------------
// state: 0 vPrev, 1 iPrev, 2..8 history
KH_HD inline void load_L(Lane& L, const int32_t* e) {
  int k = L.T(e, 2);
  L.add(e, 0, 2, 1);
  L.add(e, 1, 2, -1);
  if (L.mode == MODE_OP || L.mode == MODE_TRANOP) {
    L.add(e, 2, 0, 1);
    L.add(e, 2, 1, -1);
    return;
  }
  if (L.mode == MODE_UICINIT) {
    L.add(e, 2, 2, 1);
    L.addRHS(k, L.S(e, 1));
    return;
  }
  double Lv = L.P(e, 0);
  double req = Lv / L.H;
  double rhs = -req * L.S(e, 1);
  if (L.method == TRAP) {
    req = 2 * Lv / L.H;
    rhs = -req * L.S(e, 1) - L.S(e, 0);
  }
  L.add(e, 2, 0, 1);
  L.add(e, 2, 1, -1);
  L.add(e, 2, 2, -req);
  L.addRHS(k, rhs);
}
KH_HD inline OpPt op_L(const Lane& L, const int32_t* e, const double* x) {
  OpPt o;
  o.voltage = L.volt(e, x);
  o.current = x[L.T(e, 2)];
  o.power = o.voltage * o.current;
  o.storage = o.power;
  return o;
}
KH_HD inline void accept_L(const Lane& L, const int32_t* e, const double* x, double t, int mode) {
  if (mode == MODE_UICINIT) {
    L.S(e, 0) = 0;
    hist_reset(L, e, 2);
  } else if (mode == MODE_OP || mode == MODE_TRANOP) {
    L.S(e, 0) = 0;
    L.S(e, 1) = x[L.T(e, 2)];
    hist_reset(L, e, 2);
  } else {
    OpPt o = op_L(L, e, x);
    L.S(e, 0) = o.voltage;
    L.S(e, 1) = o.current;
  }
  hist_push(L, e, 2, t, L.P(e, 0) * L.S(e, 1));
}
KH_HD inline double trunc_L(const Lane& L, const int32_t* e, const double* x, const TruncCtx& tc) {
  OpPt o = op_L(L, e, x);
  return lte_step(L, e, 2, L.P(e, 0) * o.current, o.voltage, L.S(e, 0), tc.vnTol, tc);
}

// ---- V, I (device/source.go:102-177) -----------------------------------------
KH_HD inline void load_V(Lane& L, const int32_t* e) {
  double v = source_at(L, e);
  L.S(e, 0) = v;
  L.add(e, 0, 2, 1);
  L.add(e, 1, 2, -1);
  L.add(e, 2, 0, 1);
  L.add(e, 2, 1, -1);
  L.addRHS(L.T(e, 2), v);
}
KH_HD inline OpPt op_V(const Lane& L, const int32_t* e, const double* x) {
  OpPt o;
  o.voltage = L.volt(e, x);
  o.current = x[L.T(e, 2)];
  o.power = o.voltage * o.current;
  o.storage = 0;
  return o;
}
KH_HD inline void load_I(Lane& L, const int32_t* e) {
  double v = source_at(L, e);
  L.S(e, 0) = v;
  L.addRHS(L.T(e, 0), -v);
  L.addRHS(L.T(e, 1), v);
}
KH_HD inline OpPt op_I(const Lane& L, const int32_t* e, const double* x) {
  OpPt o;
  o.voltage = L.volt(e, x);
  o.current = L.S(e, 0);
  o.power = o.voltage * o.current;
  o.storage = 0;
  return o;
}

// ---- pnjlim (device/diode.go:214-226) --------------------------------------
KH_HD inline double pnjlim(double vnew, double vold, double vt, double vcrit, bool* lim) {
  if (vnew > vcrit && fabs(vnew - vold) > 2 * vt) {
    *lim = true;
    if (vold > 0) {
      double arg = 1 + (vnew - vold) / vt;
      if (arg > 0) return vold + vt * log(arg);
      return vcrit;
    }
    return vt * log(vnew / vt);
  }
  *lim = false;
  return vnew;
}

// ---- diode (device/diode.go:148-420); terms: a (anode), k, ai ---------------
// params: 0 IS 1 N 2 RS 3 CJO 4 VJ 5 M 6 TT 7 BV 8 IBV 9 FC 10 hasBV 11 vte 12 vcrit 13 xbv 14 grs 15 hasCharge
// state: 0 vd 1 cd 2 gd 3 iq 4 geq 5 tran 6 lim 7 gmin 8 qPrev 9 iqPrev 10 vdAcc 11 omega 12..18 history
KH_HD inline void diode_current(const Lane& L, const int32_t* e, double v, double* id, double* g) {
  double IS = L.P(e, 0), vte = L.P(e, 11);
  if (v >= -3 * vte) {
    double ex = exp(fmin(v / vte, 700.0));
    *id = IS * (ex - 1);
    *g = IS * ex / vte;
  } else if (L.P(e, 10) == 0 || v >= -L.P(e, 13)) {
    double a = 3 * vte / (v * kE);
    a = a * a * a;
    *id = -IS * (1 + a);
    *g = IS * 3 * a / v;
  } else {
    double ex = exp(fmin(-(L.P(e, 13) + v) / vte, 700.0));
    *id = -IS * ex;
    *g = IS * ex / vte;
  }
}
KH_HD inline void diode_charge(const Lane& L, const int32_t* e, double v, double id, double g, double* q, double* cc) {
  double TT = L.P(e, 6), CJO = L.P(e, 3), VJ = L.P(e, 4), M = L.P(e, 5), FC = L.P(e, 9);
  *q = TT * id;
  *cc = TT * g;
  if (CJO == 0) return;
  double fcv = FC * VJ;
  if (v < fcv) {
    double arg = 1 - v / VJ;
    double sarg = pow(arg, -M);
    *q += CJO * VJ * (1 - arg * sarg) / (1 - M);
    *cc += CJO * sarg;
    return;
  }
  double f1 = VJ * (1 - pow(1 - FC, 1 - M)) / (1 - M);
  double f2 = pow(1 - FC, 1 + M);
  double f3 = 1 - FC * (1 + M);
  *q += CJO * f1 + CJO / f2 * (f3 * (v - fcv) + M / (2 * VJ) * (v * v - fcv * fcv));
  *cc += CJO / f2 * (f3 + M * v / VJ);
}
KH_HD inline void load_D(Lane& L, const int32_t* e) {
  int k = L.T(e, 1), ai = L.T(e, 2);
  double vte = L.P(e, 11), vcrit = L.P(e, 12), xbv = L.P(e, 13);
  bool hasBV = L.P(e, 10) != 0;
  L.S(e, 7) = L.gmin;
  double vraw = L.xat(L.X, ai) - L.xat(L.X, k);
  double vd;
  bool lim;
  if (L.initJct) {
    vd = vcrit;
    lim = true;
  } else {
    double vdAcc = L.S(e, 10);
    vd = pnjlim(vraw, vdAcc, vte, vcrit, &lim);
    if (hasBV && vraw < fmin(0.0, -xbv + 10 * vte)) {
      bool l2;
      double vt = pnjlim(-(vraw + xbv), -(vdAcc + xbv), vte, vcrit, &l2);
      vd = -(vt + xbv);
      lim = l2;
    }
  }
  L.S(e, 10) = vd;
  double id, g0;
  diode_current(L, e, vd, &id, &g0);
  double cd = id + L.gmin * vd, gd = g0 + L.gmin;
  double iq = 0, geq = 0, tran = 0;
  if (L.mode == MODE_TRAN && L.P(e, 15) != 0) {
    double q, cc;
    diode_charge(L, e, vd, id, g0, &q, &cc);
    tran = 1;
    if (L.method == TRAP) {
      iq = 2 * (q - L.S(e, 8)) / L.H - L.S(e, 9);
      geq = 2 * cc / L.H;
    } else {
      iq = (q - L.S(e, 8)) / L.H;
      geq = cc / L.H;
    }
  }
  L.S(e, 0) = vd; L.S(e, 1) = cd; L.S(e, 2) = gd; L.S(e, 3) = iq; L.S(e, 4) = geq; L.S(e, 5) = tran;
  L.S(e, 6) = lim ? 1 : 0;
  double g = gd + geq;
  double ieq = cd + iq - g * vd;
  L.add(e, 2, 2, g);
  L.add(e, 1, 1, g);
  L.add(e, 2, 1, -g);
  L.add(e, 1, 2, -g);
  L.addRHS(ai, -ieq);
  L.addRHS(k, ieq);
  double grs = L.P(e, 14);
  if (grs > 0) {
    L.add(e, 0, 0, grs);
    L.add(e, 2, 2, grs);
    L.add(e, 0, 2, -grs);
    L.add(e, 2, 0, -grs);
  }
}
KH_HD inline bool conv_D(const Lane& L, const int32_t* e, const double* x, double reltol, double abstol) {
  double v = L.xat(x, L.T(e, 2)) - L.xat(x, L.T(e, 1));
  double cdhat = L.S(e, 1) + L.S(e, 2) * (v - L.S(e, 0));
  double id, g0;
  diode_current(L, e, v, &id, &g0);
  double cdn = id + L.S(e, 7) * v;
  double tol = reltol * fmax(fabs(cdhat), fabs(cdn)) + abstol;
  return fabs(cdhat - cdn) <= tol;
}
KH_HD inline OpPt op_D(const Lane& L, const int32_t* e, const double* x) {
  double v = L.xat(x, L.T(e, 2)) - L.xat(x, L.T(e, 1));
  double ir = L.S(e, 1) + L.S(e, 2) * (v - L.S(e, 0));
  double ic = L.S(e, 3) + L.S(e, 4) * (v - L.S(e, 0));
  OpPt o;
  o.current = ir + ic;
  o.voltage = L.volt(e, x);
  o.power = o.voltage * o.current;
  o.storage = v * ic;
  return o;
}
KH_HD inline void accept_D(const Lane& L, const int32_t* e, const double* x, double t, int mode) {
  double v = L.xat(x, L.T(e, 2)) - L.xat(x, L.T(e, 1));
  L.S(e, 10) = v;
  double id, g0, q, cc;
  diode_current(L, e, v, &id, &g0);
  diode_charge(L, e, v, id, g0, &q, &cc);
  if (mode == MODE_TRAN) {
    L.S(e, 9) = L.S(e, 5) != 0 ? L.S(e, 3) + L.S(e, 4) * (v - L.S(e, 0)) : 0.0;
  } else {
    L.S(e, 9) = 0;
    hist_reset(L, e, 12);
  }
  L.S(e, 8) = q;
  hist_push(L, e, 12, t, q);
}
KH_HD inline double trunc_D(const Lane& L, const int32_t* e, const double* x, const TruncCtx& tc) {
  if (L.P(e, 15) == 0) return kInf;
  double v = L.xat(x, L.T(e, 2)) - L.xat(x, L.T(e, 1));
  double id, g0, q, cc;
  diode_current(L, e, v, &id, &g0);
  diode_charge(L, e, v, id, g0, &q, &cc);
  double ic = L.S(e, 3) + L.S(e, 4) * (v - L.S(e, 0));
  return lte_step(L, e, 12, q, ic, L.S(e, 9), tc.absTol, tc);
}

// ---- MOSFET level 1 (device/mosfet.go:340-757, device/mosfet_levels.go:79-113)
// terms: 0 d, 1 s, 2 g, 3 b, 4 dp, 5 sp
// params: 0 typ 1 beta 2 vbi 3 vto 4 GAMMA 5 PHI 6 LAMBDA 7 gdpr 8 gspr 9 isd 10 iss 11 vcritD 12 vcritS
//   13 czbd 14 czbdsw 15 czbs 16 czbssw 17 PB 18 MJ 19 MJSW 20 FC 21..23 fbt 24..26 fsw 27 cgso 28 cgdo
//   29 cgbo 30 oxideCap 31 vt 32 hasJcap 33 hasGcap
enum {
  MS_VGS = 0, MS_VDS, MS_VBS, MS_VBD, MS_VGD, MS_VGB, MS_MODE, MS_CDRAIN, MS_GM, MS_GDS, MS_GMBS, MS_VON, MS_VDSAT,
  MS_CBS, MS_GBS, MS_CBD, MS_GBD, MS_LIM, MS_TRAN, MS_AG, MS_GMIN, MS_CAPG = 21, MS_CAPB = 24, MS_IQ = 26,
  MS_DV = 31, MS_AVGS = 36, MS_AVDS, MS_AVBS, MS_AVGD, MS_AVGB, MS_AVBD, MS_AVON, MS_MEYER = 43, MS_QPREV = 46,
  MS_IQPREV = 51, MS_VPREV = 56, MS_HIST = 61
};

KH_HD inline double fetlim(double vnew, double vold, double vto) {  // device/mosfet.go:347-386
  double vtsthi = fabs(2 * (vold - vto)) + 2;
  double vtstlo = vtsthi / 2 + 2;
  double vtox = vto + 3.5;
  double delv = vnew - vold;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) {
          if (-delv > vtstlo) vnew = vold - vtstlo;
        } else {
          vnew = fmax(vnew, vto + 2);
        }
      } else if (delv >= vtsthi) {
        vnew = vold + vtsthi;
      }
    } else if (delv <= 0) {
      vnew = fmax(vnew, vto - 0.5);
    } else {
      vnew = fmin(vnew, vto + 4);
    }
  } else if (delv <= 0) {
    if (-delv > vtsthi) vnew = vold - vtsthi;
  } else {
    double vtemp = vto + 0.5;
    if (vnew <= vtemp) {
      if (delv > vtstlo) vnew = vold + vtstlo;
    } else {
      vnew = vtemp;
    }
  }
  return vnew;
}

KH_HD inline double limvds(double vnew, double vold) {  // device/mosfet.go:388-404
  if (vold >= 3.5) {
    if (vnew > vold) return fmin(vnew, 3 * vold + 2);
    if (vnew < 3.5) return fmax(vnew, 2.0);
    return vnew;
  }
  if (vnew > vold) return fmin(vnew, 4.0);
  return fmax(vnew, -0.5);
}

KH_HD inline void depletion(double v, double cz, double vj, double m, double fc, double f1, double f2, double f3,
                            double* qv, double* cc) {  // device/bjt.go:358-370
  if (cz == 0) { *qv = 0; *cc = 0; return; }
  double fcv = fc * vj;
  if (v < fcv) {
    double arg = 1 - v / vj;
    double sarg = exp(-m * log(arg));
    *qv = vj * cz * (1 - arg * sarg) / (1 - m);
    *cc = cz * sarg;
    return;
  }
  double czf2 = cz / f2;
  *qv = cz * f1 + czf2 * (f3 * (v - fcv) + (m / (vj + vj)) * (v * v - fcv * fcv));
  *cc = czf2 * (f3 + m * v / vj);
}

KH_HD inline void meyer(double vgs, double vgd, double von, double vdsat, double phi, double cox, double* cgs,
                        double* cgd, double* cgb) {  // device/mosfet.go:419-438
  double vgst = vgs - von;
  vdsat = fmax(vdsat, 0.0);
  if (vgst <= -phi) { *cgs = 0; *cgd = 0; *cgb = cox; return; }
  if (vgst <= -phi / 2) { *cgs = 0; *cgd = 0; *cgb = -vgst * cox / phi; return; }
  if (vgst <= 0) { *cgs = 2 * (vgst * cox / (1.5 * phi) + cox / 3); *cgd = 0; *cgb = -vgst * cox / phi; return; }
  double vds = vgs - vgd;
  if (vdsat <= vds) { *cgs = 2 * cox / 3; *cgd = 0; *cgb = 0; return; }
  double vddif = 2 * vdsat - vds;
  double vddif1 = vdsat - vds;
  double vddif2 = vddif * vddif;
  *cgs = 2 * cox * (1 - vddif1 * vddif1 / vddif2) / 3;
  *cgd = 2 * cox * (1 - vdsat * vdsat / vddif2) / 3;
  *cgb = 0;
}

struct MosIds {
  double md, cd, gm, gds, gmbs, von, vdsat;
};

KH_HD inline MosIds mos_channel(const Lane& L, const int32_t* e, double vgs, double vds, double vbs) {
  // channel (device/mosfet.go:464-473) + the level's ids (idsLevel1,
  // device/mosfet_levels.go:79-113; levels 2 and 3 in kh_mos23.h)
  MosIds r;
  r.md = 1;
  double vg = vgs, vd = vds, vb = vbs;
  if (vds < 0) {
    r.md = -1;
    vg = vgs - vds;
    vd = -vds;
    vb = vbs - vds;
  }
  double PHI = L.P(e, 5), GAMMA = L.P(e, 4), LAMBDA = L.P(e, 6), beta = L.P(e, 1);
  if (KH_HAS(K_M23)) {  // ids, device/mosfet_levels.go:58-68: levels 2 and 3
    const int level = (int)L.P(e, 34);
    if (level == 2 || level == 3) {
      const MosLevelP lv = mos_level_params(L.par + e[ET_POFF]);
      const IdsOut o = level == 2 ? idsLevel2(lv, L.P(e, 2), beta, L.P(e, 31), PHI, GAMMA, LAMBDA, vg, vd, vb)
                                  : idsLevel3(lv, L.P(e, 2), beta, L.P(e, 31), PHI, GAMMA, L.P(e, 57), vg, vd, vb);
      r.cd = o.cd; r.gm = o.gm; r.gds = o.gds; r.gmbs = o.gmbs; r.von = o.von; r.vdsat = o.vdsat;
      return r;
    }
  }
  double sarg;
  if (vb <= 0) {
    sarg = sqrt(PHI - vb);
  } else {
    sarg = sqrt(PHI);
    sarg = fmax(0.0, sarg - vb / (sarg + sarg));
  }
  r.von = L.P(e, 2) + GAMMA * sarg;
  double vgst = vg - r.von;
  r.vdsat = fmax(vgst, 0.0);
  double arg = 0;
  if (sarg > 0) arg = GAMMA / (sarg + sarg);
  if (vgst <= 0) {
    r.cd = 0; r.gm = 0; r.gds = 0; r.gmbs = 0;
    return r;
  }
  double betap = beta * (1 + LAMBDA * vd);
  if (vgst <= vd) {
    r.cd = betap * vgst * vgst * 0.5;
    r.gm = betap * vgst;
    r.gds = LAMBDA * beta * vgst * vgst * 0.5;
  } else {
    r.cd = betap * vd * (vgst - 0.5 * vd);
    r.gm = betap * vd;
    r.gds = betap * (vgst - vd) + LAMBDA * beta * vd * (vgst - 0.5 * vd);
  }
  r.gmbs = r.gm * arg;
  return r;
}

KH_HD inline void mos_bulk_diode(const Lane& L, const int32_t* e, double v, double is, double gmin, double* i,
                                 double* g) {  // device/mosfet.go:408-415
  double vt = L.P(e, 31);
  if (v <= -3 * vt) { *i = gmin * v - is; *g = gmin; return; }
  double ex = exp(fmin(v / vt, 700.0));
  *i = is * (ex - 1) + gmin * v;
  *g = is * ex / vt + gmin;
}

KH_HD inline void mos_meyer_caps(const Lane& L, const int32_t* e, double vgs, double vgd, double von, double vdsat,
                                 double md, double out[3]) {  // device/mosfet.go:441-451
  double ox = L.P(e, 30);
  if (ox == 0) { out[0] = out[1] = out[2] = 0; return; }
  double a, b, cc;
  if (md > 0) {
    meyer(vgs, vgd, von, vdsat, L.P(e, 5), ox, &a, &b, &cc);
    out[0] = a; out[1] = b; out[2] = cc;
    return;
  }
  meyer(vgd, vgs, von, vdsat, L.P(e, 5), ox, &a, &b, &cc);
  out[0] = b; out[1] = a; out[2] = cc;
}

KH_HD inline void mos_bulk_charge(const Lane& L, const int32_t* e, double v, double cz, double czsw, double* qv,
                                  double* cc) {  // device/mosfet.go:455-460
  double q1, c1, q2, c2;
  double PB = L.P(e, 17), FC = L.P(e, 20);
  depletion(v, cz, PB, L.P(e, 18), FC, L.P(e, 21), L.P(e, 22), L.P(e, 23), &q1, &c1);
  depletion(v, czsw, PB, L.P(e, 19), FC, L.P(e, 24), L.P(e, 25), L.P(e, 26), &q2, &c2);
  *qv = q1 + q2;
  *cc = c1 + c2;
}

KH_HD inline void mos_voltages(const Lane& L, const int32_t* e, const double* x, double* vgs, double* vds, double* vbs,
                               double* vgb) {
  double t = L.P(e, 0);
  double vg = L.xat(x, L.T(e, 2)), vb = L.xat(x, L.T(e, 3)), vdp = L.xat(x, L.T(e, 4)), vsp = L.xat(x, L.T(e, 5));
  *vgs = t * (vg - vsp);
  *vds = t * (vdp - vsp);
  *vbs = t * (vb - vsp);
  *vgb = t * (vg - vb);
}

KH_HD inline double mos_cap(const Lane& L, const int32_t* e, int k) {
  return k < 3 ? L.S(e, MS_CAPG + k) : L.S(e, MS_CAPB + k - 3);
}

KH_HD inline void mos_channel_branch(const Lane& L, const int32_t* e, double* i0, double gc[3]) {  // :582-589
  if (L.S(e, MS_MODE) > 0) {
    *i0 = L.S(e, MS_CDRAIN);
    gc[0] = L.S(e, MS_GM); gc[1] = L.S(e, MS_GDS); gc[2] = L.S(e, MS_GMBS);
    return;
  }
  *i0 = -L.S(e, MS_CDRAIN);
  gc[0] = -L.S(e, MS_GM);
  gc[1] = L.S(e, MS_GM) + L.S(e, MS_GDS) + L.S(e, MS_GMBS);
  gc[2] = -L.S(e, MS_GMBS);
}

// branch, device/mosfet.go:563-578 (terminal indices into the element's terms)
KH_HD inline void mos_branch(Lane& L, const int32_t* e, int ta, int tb, double i0, int nk, const double g[3],
                             const int tp[3], const int tm[3], const double v0[3]) {
  double ieq = i0;
  for (int k = 0; k < nk; k++) {
    if (g[k] == 0) continue;
    L.add(e, ta, tp[k], g[k]);
    L.add(e, ta, tm[k], -g[k]);
    L.add(e, tb, tp[k], -g[k]);
    L.add(e, tb, tm[k], g[k]);
    ieq -= g[k] * v0[k];
  }
  ieq *= L.P(e, 0);
  L.addRHS(L.T(e, ta), -ieq);
  L.addRHS(L.T(e, tb), ieq);
}

KH_HD inline void mos_stampG(Lane& L, const int32_t* e, int ta, int tb, double g) {  // device/bjt.go:482-490
  if (g == 0) return;
  L.add(e, ta, ta, g);
  L.add(e, tb, tb, g);
  L.add(e, ta, tb, -g);
  L.add(e, tb, ta, -g);
}

KH_HD inline void load_M(Lane& L, const int32_t* e) {  // device/mosfet.go:476-551
  L.S(e, MS_GMIN) = L.gmin;
  double vgs, vds, vbs, vgb;
  mos_voltages(L, e, L.X, &vgs, &vds, &vbs, &vgb);
  bool lim = false;
  if (L.initJct) {
    vbs = -1; vgs = L.P(e, 3); vds = 0;
    lim = true;
  } else {
    double avgs = L.S(e, MS_AVGS), avds = L.S(e, MS_AVDS), avbs = L.S(e, MS_AVBS), von = L.S(e, MS_AVON);
    double vgd = vgs - vds;
    double vgdOld = avgs - avds;
    if (avds >= 0) {
      vgs = fetlim(vgs, avgs, von);
      vds = vgs - vgd;
      vds = limvds(vds, avds);
      vgd = vgs - vds;
    } else {
      vgd = fetlim(vgd, vgdOld, von);
      vds = vgs - vgd;
      vds = -limvds(-vds, -avds);
      vgs = vgd + vds;
    }
    bool l;
    if (vds >= 0) {
      vbs = pnjlim(vbs, avbs, L.P(e, 31), L.P(e, 12), &l);
    } else {
      bool l2;
      double vbd = pnjlim(vbs - vds, avbs - avds, L.P(e, 31), L.P(e, 11), &l2);
      vbs = vbd + vds;
      l = l2;
    }
    lim = l;
  }
  double vbd = vbs - vds, vgd = vgs - vds;
  vgb = vgs - vbs;
  L.S(e, MS_VGS) = vgs; L.S(e, MS_VDS) = vds; L.S(e, MS_VBS) = vbs;
  L.S(e, MS_VBD) = vbd; L.S(e, MS_VGD) = vgd; L.S(e, MS_VGB) = vgb;
  L.S(e, MS_AVGS) = vgs; L.S(e, MS_AVDS) = vds; L.S(e, MS_AVBS) = vbs;
  double cbs, gbs, cbd, gbd;
  mos_bulk_diode(L, e, vbs, L.P(e, 10), L.gmin, &cbs, &gbs);
  mos_bulk_diode(L, e, vbd, L.P(e, 9), L.gmin, &cbd, &gbd);
  L.S(e, MS_CBS) = cbs; L.S(e, MS_GBS) = gbs; L.S(e, MS_CBD) = cbd; L.S(e, MS_GBD) = gbd;
  MosIds ch = mos_channel(L, e, vgs, vds, vbs);
  L.S(e, MS_MODE) = ch.md; L.S(e, MS_CDRAIN) = ch.cd; L.S(e, MS_GM) = ch.gm; L.S(e, MS_GDS) = ch.gds;
  L.S(e, MS_GMBS) = ch.gmbs; L.S(e, MS_VON) = ch.von; L.S(e, MS_VDSAT) = ch.vdsat;
  L.S(e, MS_AVON) = ch.von;
  L.S(e, MS_LIM) = lim ? 1 : 0;
  bool tran = false;
  double ag = 0;
  double iq[5] = {0, 0, 0, 0, 0};
  if (L.mode == MODE_TRAN && (L.P(e, 33) != 0 || L.P(e, 32) != 0)) {
    tran = true;
    ag = L.method == TRAP ? 2 / L.H : 1 / L.H;
    double mc[3];
    mos_meyer_caps(L, e, vgs, vgd, ch.von, ch.vdsat, ch.md, mc);
    double ov[3] = {L.P(e, 27), L.P(e, 28), L.P(e, 29)};
    for (int k = 0; k < 3; k++) L.S(e, MS_CAPG + k) = (mc[k] + L.S(e, MS_MEYER + k)) / 2 + ov[k];
    double qbs, cbsc, qbd, cbdc;
    mos_bulk_charge(L, e, vbs, L.P(e, 15), L.P(e, 16), &qbs, &cbsc);
    mos_bulk_charge(L, e, vbd, L.P(e, 13), L.P(e, 14), &qbd, &cbdc);
    L.S(e, MS_CAPB + 0) = cbsc;
    L.S(e, MS_CAPB + 1) = cbdc;
    double qs[5] = {L.S(e, MS_QPREV + 0) + L.S(e, MS_CAPG + 0) * (vgs - L.S(e, MS_VPREV + 0)),
                    L.S(e, MS_QPREV + 1) + L.S(e, MS_CAPG + 1) * (vgd - L.S(e, MS_VPREV + 1)),
                    L.S(e, MS_QPREV + 2) + L.S(e, MS_CAPG + 2) * (vgb - L.S(e, MS_VPREV + 2)), qbs, qbd};
    for (int k = 0; k < 5; k++) {
      iq[k] = ag * (qs[k] - L.S(e, MS_QPREV + k));
      if (L.method == TRAP) iq[k] -= L.S(e, MS_IQPREV + k);
    }
  }
  L.S(e, MS_TRAN) = tran ? 1 : 0;
  L.S(e, MS_AG) = ag;
  for (int k = 0; k < 5; k++) L.S(e, MS_IQ + k) = iq[k];
  double dv[5] = {vgs, vgd, vgb, vbs, vbd};
  for (int k = 0; k < 5; k++) L.S(e, MS_DV + k) = dv[k];
  // stamp, device/mosfet.go:591-611
  double i0, gc[3];
  mos_channel_branch(L, e, &i0, gc);
  {
    int tp[3] = {2, 4, 3}, tm[3] = {5, 5, 5};
    double v0[3] = {vgs, vds, vbs};
    mos_branch(L, e, 4, 5, i0, 3, gc, tp, tm, v0);
  }
  {
    double g[3] = {gbs, 0, 0}; int tp[3] = {3, 0, 0}, tm[3] = {5, 0, 0}; double v0[3] = {vbs, 0, 0};
    mos_branch(L, e, 3, 5, cbs, 1, g, tp, tm, v0);
  }
  {
    double g[3] = {gbd, 0, 0}; int tp[3] = {3, 0, 0}, tm[3] = {4, 0, 0}; double v0[3] = {vbd, 0, 0};
    mos_branch(L, e, 3, 4, cbd, 1, g, tp, tm, v0);
  }
  if (tran) {
    const int ends[5][2] = {{2, 5}, {2, 4}, {2, 3}, {3, 5}, {3, 4}};
    for (int k = 0; k < 5; k++) {
      double g[3] = {ag * mos_cap(L, e, k), 0, 0};
      int tp[3] = {ends[k][0], 0, 0}, tm[3] = {ends[k][1], 0, 0};
      double v0[3] = {dv[k], 0, 0};
      mos_branch(L, e, ends[k][0], ends[k][1], iq[k], 1, g, tp, tm, v0);
    }
  }
  if (L.T(e, 4) != L.T(e, 0)) mos_stampG(L, e, 0, 4, L.P(e, 7));
  if (L.T(e, 5) != L.T(e, 1)) mos_stampG(L, e, 1, 5, L.P(e, 8));
}

KH_HD inline bool conv_M(const Lane& L, const int32_t* e, const double* x, double reltol, double abstol) {  // :618-631
  double vgs, vds, vbs, vgb;
  mos_voltages(L, e, x, &vgs, &vds, &vbs, &vgb);
  double vbd = vbs - vds;
  double i0, gc[3];
  mos_channel_branch(L, e, &i0, gc);
  double ich = i0 + gc[0] * (vgs - L.S(e, MS_VGS)) + gc[1] * (vds - L.S(e, MS_VDS)) + gc[2] * (vbs - L.S(e, MS_VBS));
  double ibs = L.S(e, MS_CBS) + L.S(e, MS_GBS) * (vbs - L.S(e, MS_VBS));
  double ibd = L.S(e, MS_CBD) + L.S(e, MS_GBD) * (vbd - L.S(e, MS_VBD));
  MosIds ch = mos_channel(L, e, vgs, vds, vbs);
  double ichN = ch.md * ch.cd;
  double ibsN, ibdN, gg;
  double gmin = L.S(e, MS_GMIN);
  mos_bulk_diode(L, e, vbs, L.P(e, 10), gmin, &ibsN, &gg);
  mos_bulk_diode(L, e, vbd, L.P(e, 9), gmin, &ibdN, &gg);
  double a1 = ich - ibd, b1 = ichN - ibdN, a2 = ibs + ibd, b2 = ibsN + ibdN;
  return fabs(a1 - b1) <= reltol * fmax(fabs(a1), fabs(b1)) + abstol &&
         fabs(a2 - b2) <= reltol * fmax(fabs(a2), fabs(b2)) + abstol;
}

KH_HD inline OpPt op_M(const Lane& L, const int32_t* e, const double* x) {  // :635-665
  double vgs, vds, vbs, vgb;
  mos_voltages(L, e, x, &vgs, &vds, &vbs, &vgb);
  double vbd = vbs - vds, vgd = vgs - vds;
  double t = L.P(e, 0);
  double i0, gc[3];
  mos_channel_branch(L, e, &i0, gc);
  double ich = t * (i0 + gc[0] * (vgs - L.S(e, MS_VGS)) + gc[1] * (vds - L.S(e, MS_VDS)) + gc[2] * (vbs - L.S(e, MS_VBS)));
  double ibs = t * (L.S(e, MS_CBS) + L.S(e, MS_GBS) * (vbs - L.S(e, MS_VBS)));
  double ibd = t * (L.S(e, MS_CBD) + L.S(e, MS_GBD) * (vbd - L.S(e, MS_VBD)));
  double iqv[5] = {0, 0, 0, 0, 0};
  double ps = 0;
  if (L.S(e, MS_TRAN) != 0) {
    double v[5] = {vgs, vgd, vgb, vbs, vbd};
    for (int k = 0; k < 5; k++) {
      iqv[k] = t * (L.S(e, MS_IQ + k) + L.S(e, MS_AG) * mos_cap(L, e, k) * (v[k] - L.S(e, MS_DV + k)));
      ps += v[k] * t * iqv[k];
    }
  }
  double id = ich - ibd - iqv[1] - iqv[4];
  double is = -ich - ibs - iqv[0] - iqv[3];
  double ig = iqv[0] + iqv[1] + iqv[2];
  double ib = ibs + ibd - iqv[2] + iqv[3] + iqv[4];
  double vd = L.xat(x, L.T(e, 0)), vs = L.xat(x, L.T(e, 1));
  OpPt o;
  o.voltage = vd - vs;
  o.current = id;
  o.power = vd * id + L.xat(x, L.T(e, 2)) * ig + vs * is + L.xat(x, L.T(e, 3)) * ib;
  o.storage = ps;
  return o;
}

KH_HD inline void accept_M(const Lane& L, const int32_t* e, const double* x, double t, int mode) {  // :668-705
  double vgs, vds, vbs, vgb;
  mos_voltages(L, e, x, &vgs, &vds, &vbs, &vgb);
  double vgd = vgs - vds, vbd = vbs - vds;
  MosIds ch = mos_channel(L, e, vgs, vds, vbs);
  L.S(e, MS_AVGS) = vgs; L.S(e, MS_AVDS) = vds; L.S(e, MS_AVBS) = vbs; L.S(e, MS_AVON) = ch.von;
  L.S(e, MS_AVGD) = vgd; L.S(e, MS_AVGB) = vgb; L.S(e, MS_AVBD) = vbd;
  double mc[3];
  mos_meyer_caps(L, e, vgs, vgd, ch.von, ch.vdsat, ch.md, mc);
  double v[5] = {vgs, vgd, vgb, vbs, vbd};
  double qbs, qbd, cc;
  mos_bulk_charge(L, e, vbs, L.P(e, 15), L.P(e, 16), &qbs, &cc);
  mos_bulk_charge(L, e, vbd, L.P(e, 13), L.P(e, 14), &qbd, &cc);
  if (mode == MODE_TRAN && L.S(e, MS_TRAN) != 0) {
    double qs[5];
    for (int k = 0; k < 3; k++) qs[k] = L.S(e, MS_QPREV + k) + L.S(e, MS_CAPG + k) * (v[k] - L.S(e, MS_VPREV + k));
    qs[3] = qbs;
    qs[4] = qbd;
    for (int k = 0; k < 5; k++) {
      L.S(e, MS_IQPREV + k) = L.S(e, MS_IQ + k) + L.S(e, MS_AG) * mos_cap(L, e, k) * (v[k] - L.S(e, MS_DV + k));
      L.S(e, MS_QPREV + k) = qs[k];
      hist_push(L, e, MS_HIST + 7 * k, t, qs[k]);
    }
  } else {
    double ov[3] = {L.P(e, 27), L.P(e, 28), L.P(e, 29)};
    for (int k = 0; k < 3; k++) L.S(e, MS_QPREV + k) = (mc[k] + ov[k]) * v[k];
    L.S(e, MS_QPREV + 3) = qbs;
    L.S(e, MS_QPREV + 4) = qbd;
    for (int k = 0; k < 5; k++) {
      L.S(e, MS_IQPREV + k) = 0;
      hist_reset(L, e, MS_HIST + 7 * k);
      hist_push(L, e, MS_HIST + 7 * k, t, L.S(e, MS_QPREV + k));
    }
  }
  for (int k = 0; k < 5; k++) L.S(e, MS_VPREV + k) = v[k];
  for (int k = 0; k < 3; k++) L.S(e, MS_MEYER + k) = mc[k];
}

KH_HD inline double trunc_M(const Lane& L, const int32_t* e, const double* x, const TruncCtx& tc) {  // :734-757
  if (L.S(e, MS_TRAN) == 0) return kInf;
  double vgs, vds, vbs, vgb;
  mos_voltages(L, e, x, &vgs, &vds, &vbs, &vgb);
  double v[5] = {vgs, vgs - vds, vgb, vbs, vbs - vds};
  double h = kInf;
  for (int k = 0; k < 5; k++) {
    double qk, cc;
    if (k < 3) qk = L.S(e, MS_QPREV + k) + L.S(e, MS_CAPG + k) * (v[k] - L.S(e, MS_VPREV + k));
    else if (k == 3) mos_bulk_charge(L, e, v[k], L.P(e, 15), L.P(e, 16), &qk, &cc);
    else mos_bulk_charge(L, e, v[k], L.P(e, 13), L.P(e, 14), &qk, &cc);
    if (qk == 0 && L.S(e, MS_QPREV + k) == 0) continue;
    double i = L.S(e, MS_IQ + k) + L.S(e, MS_AG) * mos_cap(L, e, k) * (v[k] - L.S(e, MS_DV + k));
    h = fmin(h, lte_step(L, e, MS_HIST + 7 * k, qk, i, L.S(e, MS_IQPREV + k), tc.absTol, tc));
  }
  return h;
}

// ---- Gummel–Poon BJT (device/bjt.go:272-765) --------------------------------
// terms: 0 c, 1 e, 2 b, 3 s (external), 4 cp, 5 bp, 6 ep (internal, or the
// external node when the series resistance is zero). The substrate junction
// runs from terminal sa to sb: (s, cp) vertical (SUBS > 0), (bp, s) lateral.
// params: the Q layout of include/kronhip.h (0 type, 1 IS, ... 51 hasCharge).
enum {
  QP_TYPE = 0, QP_IS = 1, QP_BF, QP_NF, QP_VAF, QP_IKF, QP_ISE, QP_NE, QP_BR, QP_NR, QP_VAR, QP_IKR, QP_ISC, QP_NC,
  QP_RB, QP_IRB, QP_RBM, QP_RE, QP_RC, QP_CJE, QP_VJE, QP_MJE, QP_TF, QP_XTF, QP_VTF, QP_ITF, QP_CJC, QP_VJC, QP_MJC,
  QP_XCJC, QP_TR, QP_CJS, QP_VJS, QP_MJS, QP_FC, QP_ISS, QP_NS, QP_SUBS, QP_VT, QP_VCRIT, QP_RBPR, QP_RBPI, QP_GCPR,
  QP_GEPR, QP_OVTF, QP_F1E, QP_F2E, QP_F3E, QP_F1C, QP_F2C, QP_F3C, QP_HASQ
};
// state: linearization (vbe vbc vbx vcs, the DC currents and slopes the stamp
// used, substrate current/slope, capacitances, ag, iq[4], flags, gmin),
// accepted voltages, qPrev[4], iqPrev[4], four histories.
enum {
  QS_V = 0, QS_IT = 4, QS_ITBE, QS_ITBC, QS_IBE, QS_GPI, QS_IBC, QS_GMU, QS_ISUB, QS_GSUB, QS_C = 13, QS_DQ = 17,
  QS_AG = 18, QS_IQ = 19, QS_TRAN = 23, QS_LIM = 24, QS_GMIN = 25, QS_ACC = 26, QS_QPREV = 30, QS_IQPREV = 34,
  QS_HIST = 38
};
static_assert(QS_HIST + 4 * 7 == NS_Q, "BJT state size");

KH_HD inline int bjt_sa(const Lane& L, const int32_t* e) { return L.P(e, QP_SUBS) > 0 ? 3 : 5; }
KH_HD inline int bjt_sb(const Lane& L, const int32_t* e) { return L.P(e, QP_SUBS) > 0 ? 4 : 3; }

// junction, device/bjt.go:272-283
KH_HD inline void bjt_junction(double is, double v, double nvt, double* i, double* g) {
  if (is == 0) { *i = 0; *g = 0; return; }
  if (v >= -3 * nvt) {
    double ex = exp(fmin(v / nvt, 700.0));
    *i = is * (ex - 1);
    *g = is * ex / nvt;
    return;
  }
  double a = 3 * nvt / (v * kE);
  a = a * a * a;
  *i = -is * (1 + a);
  *g = is * 3 * a / v;
}

struct BjtDC {  // bjtDC, device/bjt.go:118-127
  double cbe, gbe, cbc, gbc, it, itBE, itBC, ibe, gpi, ibc, gmu, qb, dqbdve, dqbdvc, gx;
};

// evalDC, device/bjt.go:294-355
KH_HD inline BjtDC bjt_eval_dc(const Lane& L, const int32_t* e, double vbe, double vbc, double gmin) {
  BjtDC d;
  double vt = L.P(e, QP_VT);
  bjt_junction(L.P(e, QP_IS), vbe, L.P(e, QP_NF) * vt, &d.cbe, &d.gbe);
  double cben, gben, cbcn, gbcn;
  bjt_junction(L.P(e, QP_ISE), vbe, L.P(e, QP_NE) * vt, &cben, &gben);
  cben += gmin * vbe;
  gben += gmin;
  bjt_junction(L.P(e, QP_IS), vbc, L.P(e, QP_NR) * vt, &d.cbc, &d.gbc);
  bjt_junction(L.P(e, QP_ISC), vbc, L.P(e, QP_NC) * vt, &cbcn, &gbcn);
  cbcn += gmin * vbc;
  gbcn += gmin;
  double VAF = L.P(e, QP_VAF), VAR = L.P(e, QP_VAR), IKF = L.P(e, QP_IKF), IKR = L.P(e, QP_IKR);
  double ivaf = VAF > 0 ? 1 / VAF : 0.0, ivar = VAR > 0 ? 1 / VAR : 0.0;
  double oik = IKF > 0 ? 1 / IKF : 0.0, oikr = IKR > 0 ? 1 / IKR : 0.0;
  double q1 = 1 / (1 - ivaf * vbc - ivar * vbe);
  if (oik == 0 && oikr == 0) {
    d.qb = q1;
    d.dqbdve = q1 * d.qb * ivar;
    d.dqbdvc = q1 * d.qb * ivaf;
  } else {
    double q2 = oik * d.cbe + oikr * d.cbc;
    double arg = fmax(0.0, 1 + 4 * q2);
    double sq = arg != 0 ? sqrt(arg) : 1.0;
    d.qb = q1 * (1 + sq) / 2;
    d.dqbdve = q1 * (d.qb * ivar + oik * d.gbe / sq);
    d.dqbdvc = q1 * (d.qb * ivaf + oikr * d.gbc / sq);
  }
  d.it = (d.cbe - d.cbc) / d.qb;
  d.itBE = (d.gbe - d.it * d.dqbdve) / d.qb;
  d.itBC = (-d.gbc - d.it * d.dqbdvc) / d.qb;
  d.ibe = d.cbe / L.P(e, QP_BF) + cben;
  d.gpi = d.gbe / L.P(e, QP_BF) + gben;
  d.ibc = d.cbc / L.P(e, QP_BR) + cbcn;
  d.gmu = d.gbc / L.P(e, QP_BR) + gbcn;
  double rbpr = L.P(e, QP_RBPR), rbpi = L.P(e, QP_RBPI), IRB = L.P(e, QP_IRB);
  double gx = rbpr + rbpi / d.qb;
  if (IRB > 0 && rbpi != 0) {  // current crowding
    double cb = d.ibe + d.ibc;
    double a1 = fmax(cb / IRB, 1e-9);
    double z = (-1 + sqrt(1 + 14.59025 * a1)) / 2.4317 / sqrt(a1);
    double tz = tan(z);
    gx = rbpr + 3 * rbpi * (tz - z) / z / tz / tz;
  }
  if (gx != 0) gx = 1 / gx;
  d.gx = gx;
  return d;
}

// depletion with exp(-m log arg), device/bjt.go:358-370
KH_HD inline void bjt_depletion(double v, double cz, double vj, double m, double fc, double f1, double f2, double f3,
                                double* qv, double* c) {
  if (cz == 0) { *qv = 0; *c = 0; return; }
  double fcv = fc * vj;
  if (v < fcv) {
    double arg = 1 - v / vj;
    double sarg = exp(-m * log(arg));
    *qv = vj * cz * (1 - arg * sarg) / (1 - m);
    *c = cz * sarg;
    return;
  }
  double czf2 = cz / f2;
  *qv = cz * f1 + czf2 * (f3 * (v - fcv) + (m / (vj + vj)) * (v * v - fcv * fcv));
  *c = czf2 * (f3 + m * v / vj);
}

struct BjtQ {  // bjtCharge, device/bjt.go:130-134
  double q[4], c[4], dqbeDvbc;
};

// charges, device/bjt.go:373-410
KH_HD inline BjtQ bjt_charges(const Lane& L, const int32_t* e, const double v[4], const BjtDC& d) {
  BjtQ ch;
  ch.dqbeDvbc = 0;
  double cbe = d.cbe, gbe = d.gbe;
  double TF = L.P(e, QP_TF);
  if (TF != 0 && v[0] > 0) {
    double argtf = 0, arg2 = 0, arg3 = 0, XTF = L.P(e, QP_XTF), ovtf = L.P(e, QP_OVTF), ITF = L.P(e, QP_ITF);
    if (XTF != 0) {
      argtf = XTF;
      if (ovtf != 0) argtf *= exp(v[1] * ovtf);
      arg2 = argtf;
      if (ITF != 0) {
        double t = cbe / (cbe + ITF);
        argtf *= t * t;
        arg2 = argtf * (3 - t - t);
      }
      arg3 = cbe * argtf * ovtf;
    }
    cbe = cbe * (1 + argtf) / d.qb;
    gbe = (gbe * (1 + arg2) - cbe * d.dqbdve) / d.qb;
    ch.dqbeDvbc = TF * (arg3 - cbe * d.dqbdvc) / d.qb;
  }
  double FC = L.P(e, QP_FC), qd, cdp;
  bjt_depletion(v[0], L.P(e, QP_CJE), L.P(e, QP_VJE), L.P(e, QP_MJE), FC, L.P(e, QP_F1E), L.P(e, QP_F2E),
                L.P(e, QP_F3E), &qd, &cdp);
  ch.q[0] = TF * cbe + qd;
  ch.c[0] = TF * gbe + cdp;
  double CJC = L.P(e, QP_CJC), czbc = CJC * L.P(e, QP_XCJC), VJC = L.P(e, QP_VJC), MJC = L.P(e, QP_MJC);
  double f1c = L.P(e, QP_F1C), f2c = L.P(e, QP_F2C), f3c = L.P(e, QP_F3C), TR = L.P(e, QP_TR);
  bjt_depletion(v[1], czbc, VJC, MJC, FC, f1c, f2c, f3c, &qd, &cdp);
  ch.q[1] = TR * d.cbc + qd;
  ch.c[1] = TR * d.gbc + cdp;
  bjt_depletion(v[2], CJC - czbc, VJC, MJC, FC, f1c, f2c, f3c, &ch.q[2], &ch.c[2]);
  double CJS = L.P(e, QP_CJS);
  ch.q[3] = 0;
  ch.c[3] = 0;
  if (CJS != 0) {
    double VJS = L.P(e, QP_VJS), MJS = L.P(e, QP_MJS), vcs = v[3];
    if (vcs < 0) {
      double arg = 1 - vcs / VJS;
      double sarg = exp(-MJS * log(arg));
      ch.q[3] = VJS * CJS * (1 - arg * sarg) / (1 - MJS);
      ch.c[3] = CJS * sarg;
    } else {
      ch.q[3] = vcs * CJS * (1 + MJS * vcs / (2 * VJS));
      ch.c[3] = CJS * (1 + MJS * vcs / VJS);
    }
  }
  return ch;
}

// voltages, device/bjt.go:413-418: polarity-corrected vbe, vbc, vbx, vcs.
KH_HD inline void bjt_voltages(const Lane& L, const int32_t* e, const double* x, double v[4]) {
  double t = L.P(e, QP_TYPE);
  double vb = L.xat(x, L.T(e, 2)), vbp = L.xat(x, L.T(e, 5)), vcp = L.xat(x, L.T(e, 4)), vep = L.xat(x, L.T(e, 6));
  v[0] = t * (vbp - vep);
  v[1] = t * (vbp - vcp);
  v[2] = t * (vb - vcp);
  v[3] = t * (L.xat(x, L.T(e, bjt_sa(L, e))) - L.xat(x, L.T(e, bjt_sb(L, e))));
}

// branch, device/bjt.go:447-469 (terminal indices into the element's terms):
// a current i0 + g1 dv1 + g2 dv2 from ta to tb; g2 = 0 stamps nothing more.
KH_HD inline void bjt_branch(Lane& L, const int32_t* e, int ta, int tb, double i0, double g1, int p1, int n1,
                             double v10, double g2, int p2, int n2, double v20) {
  if (g1 != 0) {
    L.add(e, ta, p1, g1); L.add(e, ta, n1, -g1); L.add(e, tb, p1, -g1); L.add(e, tb, n1, g1);
  }
  if (g2 != 0) {
    L.add(e, ta, p2, g2); L.add(e, ta, n2, -g2); L.add(e, tb, p2, -g2); L.add(e, tb, n2, g2);
  }
  double ieq = L.P(e, QP_TYPE) * (i0 - g1 * v10 - g2 * v20);
  L.addRHS(L.T(e, ta), -ieq);
  L.addRHS(L.T(e, tb), ieq);
}

// Load + stamp, device/bjt.go:421-515
KH_HD inline void load_Q(Lane& L, const int32_t* e) {
  double gmin = L.gmin;
  L.S(e, QS_GMIN) = gmin;
  double v[4];
  bjt_voltages(L, e, L.X, v);
  bool lim = false;
  if (L.initJct) {
    v[0] = L.P(e, QP_VCRIT); v[1] = 0; v[2] = 0; v[3] = 0;
    lim = true;
  } else {
    bool l1, l2;
    double vt = L.P(e, QP_VT), vcrit = L.P(e, QP_VCRIT);
    v[0] = pnjlim(v[0], L.S(e, QS_ACC + 0), vt, vcrit, &l1);
    v[1] = pnjlim(v[1], L.S(e, QS_ACC + 1), vt, vcrit, &l2);
    lim = l1 || l2;
  }
  L.S(e, QS_ACC + 0) = v[0];
  L.S(e, QS_ACC + 1) = v[1];
  for (int k = 0; k < 4; k++) L.S(e, QS_V + k) = v[k];
  BjtDC d = bjt_eval_dc(L, e, v[0], v[1], gmin);
  L.S(e, QS_IT) = d.it; L.S(e, QS_ITBE) = d.itBE; L.S(e, QS_ITBC) = d.itBC; L.S(e, QS_IBE) = d.ibe;
  L.S(e, QS_GPI) = d.gpi; L.S(e, QS_IBC) = d.ibc; L.S(e, QS_GMU) = d.gmu;
  double isub, gsub;  // substrate, device/bjt.go:288-291
  bjt_junction(L.P(e, QP_ISS), v[3], L.P(e, QP_NS) * L.P(e, QP_VT), &isub, &gsub);
  isub += gmin * v[3];
  gsub += gmin;
  L.S(e, QS_ISUB) = isub;
  L.S(e, QS_GSUB) = gsub;
  bool tran = false;
  double ag = 0, iq[4] = {0, 0, 0, 0};
  BjtQ ch;
  if (L.mode == MODE_TRAN && L.P(e, QP_HASQ) != 0) {
    ch = bjt_charges(L, e, v, d);
    tran = true;
    if (L.method == TRAP) {
      ag = 2 / L.H;
      for (int k = 0; k < 4; k++) iq[k] = ag * (ch.q[k] - L.S(e, QS_QPREV + k)) - L.S(e, QS_IQPREV + k);
    } else {
      ag = 1 / L.H;
      for (int k = 0; k < 4; k++) iq[k] = ag * (ch.q[k] - L.S(e, QS_QPREV + k));
    }
    for (int k = 0; k < 4; k++) L.S(e, QS_C + k) = ch.c[k];
    L.S(e, QS_DQ) = ch.dqbeDvbc;
  }
  L.S(e, QS_AG) = ag;
  for (int k = 0; k < 4; k++) L.S(e, QS_IQ + k) = iq[k];
  L.S(e, QS_TRAN) = tran ? 1 : 0;
  L.S(e, QS_LIM) = lim ? 1 : 0;
  const int sa = bjt_sa(L, e), sb = bjt_sb(L, e);
  bjt_branch(L, e, 4, 6, d.it, d.itBE, 5, 6, v[0], d.itBC, 5, 4, v[1]);
  bjt_branch(L, e, 5, 6, d.ibe, d.gpi, 5, 6, v[0], 0, 0, 0, 0);
  bjt_branch(L, e, 5, 4, d.ibc, d.gmu, 5, 4, v[1], 0, 0, 0, 0);
  bjt_branch(L, e, sa, sb, isub, gsub, sa, sb, v[3], 0, 0, 0, 0);
  if (tran) {
    bjt_branch(L, e, 5, 6, iq[0], ag * ch.c[0], 5, 6, v[0], ag * ch.dqbeDvbc, 5, 4, v[1]);
    bjt_branch(L, e, 5, 4, iq[1], ag * ch.c[1], 5, 4, v[1], 0, 0, 0, 0);
    bjt_branch(L, e, 2, 4, iq[2], ag * ch.c[2], 2, 4, v[2], 0, 0, 0, 0);
    bjt_branch(L, e, sa, sb, iq[3], ag * ch.c[3], sa, sb, v[3], 0, 0, 0, 0);
  }
  if (L.T(e, 5) != L.T(e, 2)) mos_stampG(L, e, 2, 5, d.gx);
  if (L.T(e, 4) != L.T(e, 0)) mos_stampG(L, e, 0, 4, L.P(e, QP_GCPR));
  if (L.T(e, 6) != L.T(e, 1)) mos_stampG(L, e, 1, 6, L.P(e, QP_GEPR));
}

// ConvTest, device/bjt.go:521-532
KH_HD inline bool conv_Q(const Lane& L, const int32_t* e, const double* x, double reltol, double abstol) {
  double v[4];
  bjt_voltages(L, e, x, v);
  double dbe = v[0] - L.S(e, QS_V + 0), dbc = v[1] - L.S(e, QS_V + 1);
  double it = L.S(e, QS_IT), ibc = L.S(e, QS_IBC), gmu = L.S(e, QS_GMU);
  double ichat = it - ibc + L.S(e, QS_ITBE) * dbe + (L.S(e, QS_ITBC) - gmu) * dbc;
  double ibhat = L.S(e, QS_IBE) + ibc + L.S(e, QS_GPI) * dbe + gmu * dbc;
  BjtDC n = bjt_eval_dc(L, e, v[0], v[1], L.S(e, QS_GMIN));
  double ic = n.it - n.ibc, ib = n.ibe + n.ibc;
  return fabs(ichat - ic) <= reltol * fmax(fabs(ichat), fabs(ic)) + abstol &&
         fabs(ibhat - ib) <= reltol * fmax(fabs(ibhat), fabs(ib)) + abstol;
}

// terminalCurrents + Operating, device/bjt.go:537-579
KH_HD inline OpPt op_Q(const Lane& L, const int32_t* e, const double* x) {
  double v[4];
  bjt_voltages(L, e, x, v);
  double t = L.P(e, QP_TYPE);
  double dbe = v[0] - L.S(e, QS_V + 0), dbc = v[1] - L.S(e, QS_V + 1);
  double it = t * (L.S(e, QS_IT) + L.S(e, QS_ITBE) * dbe + L.S(e, QS_ITBC) * dbc);
  double ibe = t * (L.S(e, QS_IBE) + L.S(e, QS_GPI) * dbe);
  double ibc = t * (L.S(e, QS_IBC) + L.S(e, QS_GMU) * dbc);
  double iqbe = 0, iqbc = 0, iqbx = 0, iqcs = 0, ps = 0;
  if (L.S(e, QS_TRAN) != 0) {
    double ag = L.S(e, QS_AG);
    iqbe = t * (L.S(e, QS_IQ + 0) + ag * L.S(e, QS_C + 0) * dbe + ag * L.S(e, QS_DQ) * dbc);
    iqbc = t * (L.S(e, QS_IQ + 1) + ag * L.S(e, QS_C + 1) * dbc);
    iqbx = t * (L.S(e, QS_IQ + 2) + ag * L.S(e, QS_C + 2) * (v[2] - L.S(e, QS_V + 2)));
    iqcs = t * (L.S(e, QS_IQ + 3) + ag * L.S(e, QS_C + 3) * (v[3] - L.S(e, QS_V + 3)));
    ps = v[0] * t * iqbe + v[1] * t * iqbc + v[2] * t * iqbx + v[3] * t * iqcs;
  }
  double isb = t * (L.S(e, QS_ISUB) + L.S(e, QS_GSUB) * (v[3] - L.S(e, QS_V + 3))) + iqcs;
  double ic = it - ibc - iqbc - iqbx;
  double ib = ibe + ibc + iqbe + iqbc + iqbx;
  double ie = -it - ibe - iqbe;
  double is;
  if (L.P(e, QP_SUBS) > 0) {
    is = isb;
    ic = ic - isb;
  } else {
    ib = ib + isb;
    is = -isb;
  }
  double vc = L.xat(x, L.T(e, 0)), ve = L.xat(x, L.T(e, 1));
  OpPt o;
  o.power = vc * ic + L.xat(x, L.T(e, 2)) * ib + ve * ie + L.xat(x, L.T(e, 3)) * is;
  o.voltage = vc - ve;
  o.current = ic;
  o.storage = ps;
  return o;
}

// Accept, device/bjt.go:582-605
KH_HD inline void accept_Q(const Lane& L, const int32_t* e, const double* x, double t, int mode) {
  double v[4];
  bjt_voltages(L, e, x, v);
  BjtDC d = bjt_eval_dc(L, e, v[0], v[1], L.S(e, QS_GMIN));
  BjtQ ch = bjt_charges(L, e, v, d);
  bool tr = mode == MODE_TRAN && L.S(e, QS_TRAN) != 0;
  double ag = L.S(e, QS_AG);
  for (int k = 0; k < 4; k++) {
    double i = 0;
    if (tr) {
      i = L.S(e, QS_IQ + k) + ag * L.S(e, QS_C + k) * (v[k] - L.S(e, QS_V + k));
      if (k == 0) i += ag * L.S(e, QS_DQ) * (v[1] - L.S(e, QS_V + 1));
    }
    L.S(e, QS_IQPREV + k) = i;
    if (mode != MODE_TRAN) hist_reset(L, e, QS_HIST + 7 * k);
    L.S(e, QS_QPREV + k) = ch.q[k];
    hist_push(L, e, QS_HIST + 7 * k, t, ch.q[k]);
  }
  for (int k = 0; k < 4; k++) L.S(e, QS_ACC + k) = v[k];
}

// Truncate, device/bjt.go:655-675
KH_HD inline double trunc_Q(const Lane& L, const int32_t* e, const double* x, const TruncCtx& tc) {
  if (L.P(e, QP_HASQ) == 0 || L.S(e, QS_TRAN) == 0) return kInf;
  double v[4];
  bjt_voltages(L, e, x, v);
  BjtDC d = bjt_eval_dc(L, e, v[0], v[1], L.S(e, QS_GMIN));
  BjtQ ch = bjt_charges(L, e, v, d);
  double ag = L.S(e, QS_AG);
  double h = kInf;
  for (int k = 0; k < 4; k++) {
    if (ch.q[k] == 0 && L.S(e, QS_QPREV + k) == 0) continue;
    double i = L.S(e, QS_IQ + k) + ag * L.S(e, QS_C + k) * (v[k] - L.S(e, QS_V + k));
    if (k == 0) i += ag * L.S(e, QS_DQ) * (v[1] - L.S(e, QS_V + 1));
    h = fmin(h, lte_step(L, e, QS_HIST + 7 * k, ch.q[k], i, L.S(e, QS_IQPREV + k), tc.absTol, tc));
  }
  return h;
}

// device.Operating of any element (the resistor's, device/passive.go:88-97,
// without spin-connection terms).
KH_HD inline OpPt operating_point(const Lane& L, const int32_t* e, const double* x) {
  switch (e[ET_KIND]) {
    case K_R: {
      OpPt o;
      o.voltage = L.volt(e, x);
      o.current = o.voltage * L.P(e, 1);
      o.power = o.voltage * o.current;
      o.storage = 0;
      return o;
    }
    case K_C: return op_C(L, e, x);
    case K_L: return op_L(L, e, x);
    case K_V: return op_V(L, e, x);
    case K_I: return op_I(L, e, x);
    case K_D: if (KH_HAS(K_D)) return op_D(L, e, x); break;
    case K_M: if (KH_HAS(K_M)) return op_M(L, e, x); break;
    case K_Q: if (KH_HAS(K_Q)) return op_Q(L, e, x); break;
    case K_E: case K_F: case K_G: case K_H: return op_ctl(L, e, x);
    case K_K: return op_K();
  }
  OpPt o = {0, 0, 0, 0};
  return o;
}

}  // namespace kh
