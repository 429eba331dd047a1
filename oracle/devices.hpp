// ORACLE — test infrastructure only (see sparse.hpp).
//
// CPU restatement of kronspice's device models for the elements the solve path
// uses: R (device/passive.go), C and L (device/reactive.go), V and I
// (device/source.go, device/waveform.go), the SPICE3 diode (device/diode.go),
// the level-1 MOSFET (device/mosfet.go, device/mosfet_levels.go), the
// Gummel–Poon BJT (device/bjt.go) and the LTE
// estimate (device/lte.go). Parameters arrive already resolved by the host
// setup step in the layout of include/kronhip.h.
#pragma once
#include <cmath>
#include <complex>
#include <limits>
#include <vector>
#include "../include/kronhip.h"

namespace oracle {

using cplx = std::complex<double>;
constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kE = 2.71828182845904523536028747135266249775724709369995957496696763;  // math.E
constexpr double kPi = 3.14159265358979323846264338327950288419716939937510582097494459;  // math.Pi

enum Mode { ModeOP = 0, ModeTranOP = 1, ModeUICInit = 2, ModeTran = 3 };  // device/device.go:105-116
enum Method { BackwardEuler = 1, Trapezoidal = 2 };                       // device/device.go:121-124

struct Stamper {
  virtual void add(int i, int j, double v) = 0;
  virtual ~Stamper() {}
};
struct CStamper {
  virtual void add(int i, int j, cplx v) = 0;
  virtual ~CStamper() {}
};

// LoadContext, device/device.go:127-148
struct LoadContext {
  Stamper* M = nullptr;
  std::vector<double>* RHS = nullptr;
  int mode = ModeOP;
  double time = 0, H = 0;
  int method = Trapezoidal;
  const double* X = nullptr;
  int nx = 0;
  bool initJct = false;
  double gmin = 1e-12;
  bool sourceStepping = false;
  double srcFact = 0;
  void addRHS(int i, double v) {  // :151-155
    if (i >= 0 && i < (int)RHS->size()) (*RHS)[i] += v;
  }
};

struct ACContext {  // device/device.go:162-174
  CStamper* M = nullptr;
  std::vector<cplx>* RHS = nullptr;
  double omega = 0;
  void addRHS(int i, cplx v) {
    if (i >= 0 && i < (int)RHS->size()) (*RHS)[i] += v;
  }
};

struct TruncContext {  // device/lte.go:10-19
  double time = 0, H = 0;
  int method = Trapezoidal;
  double relTol = 1e-3, absTol = 1e-12, vnTol = 1e-6, chgTol = 1e-14, trTol = 7;
};

// history, device/lte.go:23-37
struct History {
  double t[3] = {0, 0, 0}, q[3] = {0, 0, 0};
  int n = 0;
  void push(double tt, double qq) {
    t[2] = t[1]; t[1] = t[0]; t[0] = tt;
    q[2] = q[1]; q[1] = q[0]; q[0] = qq;
    if (n < 3) n++;
  }
  void reset() { n = 0; }
};

// lteStep, device/lte.go:50-85
inline double lteStep(const History& hist, double qn, double in, double inPrev, double absTol, const TruncContext& tc) {
  int k = tc.method;
  if (k < 1 || k > 2 || hist.n < k + 1) return kInf;
  double ts[4] = {tc.time, hist.t[0], hist.t[1], hist.t[2]};
  double qs[4] = {qn, hist.q[0], hist.q[1], hist.q[2]};
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
  double fact = 2.0, c = 0.5;
  if (k == 2) { fact = 6; c = 1.0 / 12; }
  double h = tc.H;
  // Go's math.Pow(x, 1) is x, Pow(x, 0.5) is Sqrt(x) and Pow(x, 2) squares
  // the mantissa once then scales by a power of two (math/pow.go): x*x.
  double errI = c * (k == 1 ? h : h * h) * std::fabs(deriv * fact);
  if (errI == 0 || std::isnan(errI)) return kInf;
  double qmax = std::max(std::fabs(qn), std::fabs(hist.q[0]));
  double allowed = tc.trTol * std::max(tc.relTol * std::max(std::fabs(in), std::fabs(inPrev)) + absTol,
                                       tc.relTol * std::max(qmax, tc.chgTol) / h);
  return h * (k == 1 ? allowed / errI : std::sqrt(allowed / errI));
}

inline double at(const double* x, int nx, int i) { return (i < 0 || i >= nx) ? 0.0 : x[i]; }  // device/diode.go:140-145

struct OperatingPoint {  // device/device.go:186-197
  double voltage = 0, current = 0, power = 0, storage = 0;
};

// Ledger accumulator, em/energy.go:25-79
struct Accumulator {
  double inflow = 0, outflow = 0, dissipation = 0, storageRate = 0, modelSupplied = 0, scale = 0;
  // channel: 0 electrical, 2 thermal (exchange/channel.go:28-31)
  int lastCh = -1;     // the channel and power of the latest addPort (one element's probe)
  double lastPw = 0;
  void addPort(int ch, double pw) {
    lastCh = ch;
    lastPw = pw;
    if (pw == 0) return;
    scale += std::fabs(pw);
    if (ch == 2 && pw < 0) dissipation += -pw;
    else if (pw > 0) inflow += pw;
    else outflow += -pw;
  }
  void addStorage(double p) { storageRate += p; scale += std::fabs(p); }
  void addModelSupplied(double p) {
    if (p == 0) return;
    modelSupplied += p;
    scale += std::fabs(p);
  }
  double residual() const { return inflow + modelSupplied - outflow - dissipation - storageRate; }
};

struct SetupContext {  // device/device.go:36-100
  int numNodes = 1, branches = 0;
  std::vector<int> internal;
  int unknown(int node) const { return node - 1; }
  int allocBranch() { return numNodes - 1 + branches++; }
  int allocUnknown() { int i = allocBranch(); internal.push_back(i); return i; }
  int size() const { return numNodes - 1 + branches; }
  std::vector<char> voltageUnknowns() const {
    std::vector<char> v(size(), 0);
    for (int i = 0; i < numNodes - 1; i++) v[i] = 1;
    for (int i : internal) v[i] = 1;
    return v;
  }
};

// Resolved source waveform, device/waveform.go:16-239 (PULSE, SIN, EXP).
struct Wave {
  int kind = 0;  // 0 none, 1 pulse, 2 sin, 3 exp
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  double value(double t) const {  // :69-126
    switch (kind) {
      case 1: {
        double v1 = a[0], v2 = a[1], td = a[2], tr = a[3], tf = a[4], pw = a[5], per = a[6];
        if (t < td) return v1;
        double tt = t - td;
        if (per > 0) tt = std::fmod(tt, per);
        if (tt < tr) return v1 + (v2 - v1) * tt / tr;
        if (tt < tr + pw) return v2;
        if (tt < tr + pw + tf) return v2 + (v1 - v2) * (tt - tr - pw) / tf;
        return v1;
      }
      case 2: {
        double vo = a[0], va = a[1], f = a[2], td = a[3], theta = a[4], ph = a[5] * kPi / 180;
        if (t < td) return vo + va * std::sin(ph);
        double tt = t - td;
        return vo + va * std::exp(-tt * theta) * std::sin(2 * kPi * f * tt + ph);
      }
      case 3: {
        double v1 = a[0], v2 = a[1], td1 = a[2], tau1 = a[3], td2 = a[4], tau2 = a[5];
        if (t < td1) return v1;
        if (t < td2) return v1 + (v2 - v1) * (1 - std::exp(-(t - td1) / tau1));
        return v1 + (v2 - v1) * (1 - std::exp(-(t - td1) / tau1)) + (v1 - v2) * (1 - std::exp(-(t - td2) / tau2));
      }
    }
    return 0;
  }
  double nextBreakpoint(double t) const {  // :190-239
    const double rel = 1e-12;
    auto after = [&](double b) { return b > t + rel * std::max(std::fabs(t), 1e-15); };
    switch (kind) {
      case 1: {
        double td = a[2], tr = a[3], tf = a[4], pw = a[5], per = a[6];
        double edges[4] = {0, tr, tr + pw, tr + pw + tf};
        if (per <= 0) {
          for (double e : edges)
            if (after(td + e)) return td + e;
          return kInf;
        }
        double k = std::floor((t - td) / per);
        if (k < 0) k = 0;
        for (double kk = k; kk <= k + 1; kk++) {
          double base = td + kk * per;
          for (double e : edges)
            if (after(base + e)) return base + e;
        }
        return td + (k + 2) * per;
      }
      case 2:
        if (after(a[3])) return a[3];
        break;
      case 3:
        if (after(a[2])) return a[2];
        if (after(a[4])) return a[4];
        break;
    }
    return kInf;
  }
};

// integrate, device/bjt.go:603-622: 16-point Gauss-Legendre (Abramowitz &
// Stegun Table 25.4, the 16 digits the reference carries) on n equal panels.
static const double kGLx[8] = {0.0950125098376374, 0.2816035507792589, 0.4580167776572274, 0.6178762444026438,
                               0.7554044083550030, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499};
static const double kGLw[8] = {0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                               0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541};
template <class F>
static double glIntegrate(const F& f, double a, double b, int n) {
  double h = (b - a) / n, s = 0.0;
  for (int p = 0; p < n; p++) {
    double mid = a + (p + 0.5) * h, half = h / 2;
    for (int k = 0; k < 8; k++) s += kGLw[k] * (f(mid - half * kGLx[k]) + f(mid + half * kGLx[k]));
  }
  return s * h / 2;
}

struct Device {
  int kind = 0;
  int nodes[4] = {0, 0, 0, 0};
  int u[2] = {-1, -1};
  virtual ~Device() {}
  void setupNodes(SetupContext& c) { u[0] = c.unknown(nodes[0]); u[1] = c.unknown(nodes[1]); }
  double voltage(const double* x) const {  // device/device.go:310-319
    double v = 0;
    if (u[0] >= 0) v += x[u[0]];
    if (u[1] >= 0) v -= x[u[1]];
    return v;
  }
  cplx voltageAC(const cplx* x) const {
    cplx v = 0;
    if (u[0] >= 0) v += x[u[0]];
    if (u[1] >= 0) v -= x[u[1]];
    return v;
  }
  virtual void setup(SetupContext& c) { setupNodes(c); }
  virtual void load(LoadContext& lc) = 0;
  virtual OperatingPoint operating(const double* x, int nx) = 0;
  virtual void accept(const double*, int, double, int) {}
  virtual bool reactive() const { return false; }
  virtual double truncate(const double*, int, const TruncContext&) { return kInf; }
  virtual bool nonlinear() const { return false; }
  virtual bool limited() const { return false; }
  virtual bool convTest(const double*, int, double, double) { return true; }
  virtual double nextBreakpoint(double) const { return kInf; }
  virtual bool breakpointer() const { return false; }
  // Ledger rows in channel order; returns thermal/electrical powers.
  virtual void ports(const OperatingPoint& op, Accumulator& acc) = 0;
  virtual bool modelSupplier() const { return false; }
  virtual double modelSupplied(const OperatingPoint&) const { return 0; }
  // device.Reactive.StoredEnergy at solution x (0 for elements that store nothing)
  virtual double storedEnergy(const double*, int) { return 0; }
  virtual void loadAC(ACContext& ac) = 0;
  virtual void initialCondition() {}
};

inline void stampY(CStamper* m, int a, int b, cplx y) {  // device/device.go:333-338
  m->add(a, a, y);
  m->add(b, b, y);
  m->add(a, b, -y);
  m->add(b, a, -y);
}

// Resistor, device/passive.go:32-134 (no ECE terms).
struct Resistor : Device {
  double R = 0, g = 0;
  void load(LoadContext& lc) override {
    int a = u[0], b = u[1];
    lc.M->add(a, a, g);
    lc.M->add(b, b, g);
    lc.M->add(a, b, -g);
    lc.M->add(b, a, -g);
  }
  OperatingPoint operating(const double* x, int) override {
    OperatingPoint op;
    op.voltage = voltage(x);
    op.current = op.voltage * g;
    op.power = op.voltage * op.current;
    return op;
  }
  void ports(const OperatingPoint& op, Accumulator& acc) override {
    if (R > 0) acc.addPort(2, -(op.power));  // ExteriorPower, :99-104
  }
  bool modelSupplier() const override { return true; }
  double modelSupplied(const OperatingPoint& op) const override { return R < 0 ? -op.power : 0; }
  void loadAC(ACContext& ac) override { stampY(ac.M, u[0], u[1], cplx(g, 0)); }
};

// Capacitor, device/reactive.go:25-184, :355-374
struct Capacitor : Device {
  double C = 0, ic = 0;
  bool hasIC = false;
  double geq = 0, ihist = 0;
  bool tran = false;
  double omega = 0;
  double vPrev = 0, iPrev = 0;
  History hist;
  void load(LoadContext& lc) override {
    int a = u[0], b = u[1];
    if (lc.mode == ModeOP || lc.mode == ModeTranOP) {
      tran = false; geq = 0; ihist = 0;
      return;
    }
    if (lc.mode == ModeUICInit) {
      tran = false;
      double g = 1e9;  // uicConductance, :57
      lc.M->add(a, a, g); lc.M->add(b, b, g); lc.M->add(a, b, -g); lc.M->add(b, a, -g);
      lc.addRHS(a, g * vPrev);
      lc.addRHS(b, -g * vPrev);
      geq = 0; ihist = 0;
      return;
    }
    tran = true;
    if (lc.method == Trapezoidal) {
      geq = 2 * C / lc.H;
      ihist = -(geq * vPrev + iPrev);
    } else {
      geq = C / lc.H;
      ihist = -geq * vPrev;
    }
    double g = geq;
    lc.M->add(a, a, g); lc.M->add(b, b, g); lc.M->add(a, b, -g); lc.M->add(b, a, -g);
    lc.addRHS(a, -ihist);
    lc.addRHS(b, ihist);
  }
  OperatingPoint operating(const double* x, int) override {
    OperatingPoint op;
    op.voltage = voltage(x);
    double i = 0;
    if (tran) i = geq * op.voltage + ihist;
    op.current = i;
    op.power = op.voltage * i;
    op.storage = op.voltage * i;
    return op;
  }
  void accept(const double* x, int nx, double t, int mode) override {
    if (mode == ModeUICInit) {
      iPrev = 0;
      hist.reset();
    } else if (mode == ModeOP || mode == ModeTranOP) {
      vPrev = voltage(x);
      iPrev = 0;
      hist.reset();
    } else {
      OperatingPoint op = operating(x, nx);
      vPrev = op.voltage;
      iPrev = op.current;
    }
    hist.push(t, C * vPrev);
  }
  void initialCondition() override { vPrev = hasIC ? ic : 0; iPrev = 0; }
  bool reactive() const override { return true; }
  double truncate(const double* x, int nx, const TruncContext& tc) override {
    OperatingPoint op = operating(x, nx);
    return lteStep(hist, C * op.voltage, op.current, iPrev, tc.absTol, tc);
  }
  double storedEnergy(const double* x, int nx) override {  // StoredEnergy, :171-175
    double v = at(x, nx, u[0]) - at(x, nx, u[1]);
    return 0.5 * C * v * v;
  }
  void ports(const OperatingPoint&, Accumulator&) override {}
  void loadAC(ACContext& ac) override {
    omega = ac.omega;
    stampY(ac.M, u[0], u[1], cplx(0, ac.omega * C));
  }
};

// Inductor, device/reactive.go:199-352, :376-401
struct Inductor : Device {
  double L = 0, ic = 0;
  bool hasIC = false;
  int branch = -1;
  double vPrev = 0, iPrev = 0;
  History hist;
  void setup(SetupContext& c) override {
    setupNodes(c);
    branch = c.allocBranch();
  }
  void load(LoadContext& lc) override {
    int a = u[0], b = u[1], k = branch;
    lc.M->add(a, k, 1);
    lc.M->add(b, k, -1);
    if (lc.mode == ModeOP || lc.mode == ModeTranOP) {
      lc.M->add(k, a, 1);
      lc.M->add(k, b, -1);
      return;
    }
    if (lc.mode == ModeUICInit) {
      lc.M->add(k, k, 1);
      lc.addRHS(k, iPrev);
      return;
    }
    double req = L / lc.H;
    double rhs = -req * iPrev;
    if (lc.method == Trapezoidal) {
      req = 2 * L / lc.H;
      rhs = -req * iPrev - vPrev;
    }
    lc.M->add(k, a, 1);
    lc.M->add(k, b, -1);
    lc.M->add(k, k, -req);
    lc.addRHS(k, rhs);
  }
  OperatingPoint operating(const double* x, int) override {
    OperatingPoint op;
    op.voltage = voltage(x);
    op.current = x[branch];
    op.power = op.voltage * op.current;
    op.storage = op.power;
    return op;
  }
  void accept(const double* x, int nx, double t, int mode) override {
    if (mode == ModeUICInit) {
      vPrev = 0;
      hist.reset();
    } else if (mode == ModeOP || mode == ModeTranOP) {
      vPrev = 0;
      iPrev = x[branch];
      hist.reset();
    } else {
      OperatingPoint op = operating(x, nx);
      vPrev = op.voltage;
      iPrev = op.current;
    }
    hist.push(t, L * iPrev);
  }
  void initialCondition() override { vPrev = 0; iPrev = hasIC ? ic : 0; }
  bool reactive() const override { return true; }
  double truncate(const double* x, int nx, const TruncContext& tc) override {
    OperatingPoint op = operating(x, nx);
    return lteStep(hist, L * op.current, op.voltage, vPrev, tc.vnTol, tc);
  }
  double storedEnergy(const double* x, int) override {  // StoredEnergy, :339-342
    double i = x[branch];
    return 0.5 * L * i * i;
  }
  void ports(const OperatingPoint&, Accumulator&) override {}
  void loadAC(ACContext& ac) override {
    int a = u[0], b = u[1], k = branch;
    ac.M->add(a, k, 1);
    ac.M->add(b, k, -1);
    ac.M->add(k, a, 1);
    ac.M->add(k, b, -1);
    ac.M->add(k, k, cplx(0, -ac.omega * L));
  }
};

// Mutual inductance (SPICE K), device/mutual.go:13-157 without the
// spin-connection terms: couples the branch rows of two inductors (k1, k2,
// found by Link); nothing at DC; in a transient step each inductor's
// integration formula applied to the flux the other's current links; -jwM in
// AC. Stores M·i1·i2, has no port and no voltage or current of its own.
struct Mutual : Device {
  double M = 0;
  int l[2] = {-1, -1};  // the coupled inductors' element indices
  int k1 = -1, k2 = -1;
  double i1Prev = 0, i2Prev = 0;
  void setup(SetupContext&) override {}  // Setup allocates nothing (:54-55)
  void load(LoadContext& lc) override {  // :74-113
    if (lc.mode == ModeOP || lc.mode == ModeTranOP || lc.mode == ModeUICInit) return;
    double f = M / lc.H;
    if (lc.method == Trapezoidal) f = 2 * M / lc.H;
    lc.M->add(k1, k2, -f);
    lc.M->add(k2, k1, -f);
    lc.addRHS(k1, -f * i2Prev);
    lc.addRHS(k2, -f * i1Prev);
  }
  OperatingPoint operating(const double*, int) override { return OperatingPoint(); }  // :115-126
  void accept(const double* x, int, double, int) override {  // :128-131
    i1Prev = x[k1];
    i2Prev = x[k2];
  }
  bool reactive() const override { return true; }  // Truncate: +Inf (:135-136)
  double storedEnergy(const double* x, int) override { return M * x[k1] * x[k2]; }  // :132-133
  void ports(const OperatingPoint&, Accumulator&) override {}
  void loadAC(ACContext& ac) override {  // :141-148
    const cplx z(-0.0, -ac.omega * M);
    ac.M->add(k1, k2, z - cplx(0, 0));
    ac.M->add(k2, k1, z + cplx(0, 0));
  }
};

// sourceValue, device/source.go:21-73
struct SourceValue {
  double dc = 0;
  bool hasDC = false;
  Wave w;
  cplx ac = 0;
  double override_ = 0;  // DC value forced by a .DC sweep (SetDCOverride, device/source.go:179-203)
  bool overridden = false;
  double at(const LoadContext& lc) const {
    double v = raw(lc);
    if (lc.sourceStepping && lc.mode != ModeTran) v *= lc.srcFact;
    return v;
  }
  double raw(const LoadContext& lc) const {
    if (overridden && lc.mode == ModeOP) return override_;
    if (w.kind == 0) return dc;
    if (lc.mode == ModeOP && hasDC) return dc;
    double t = lc.time;
    if (lc.mode != ModeTran) t = 0;
    return w.value(t);
  }
  double nextBreakpoint(double t) const { return w.kind == 0 ? kInf : w.nextBreakpoint(t); }
};

inline void loadSource(SourceValue& s, const double* p) {
  s.dc = p[0];
  s.hasDC = p[1] != 0;
  s.w.kind = (int)p[2];
  for (int k = 0; k < 7; k++) s.w.a[k] = p[3 + k];
  double mag = p[10], ph = p[11] * kPi / 180;
  s.ac = cplx(mag * std::cos(ph), mag * std::sin(ph));  // cmplx.Rect
}

// VoltageSource, device/source.go:77-129, :186-199
struct VoltageSource : Device {
  SourceValue src;
  int branch = -1;
  double value = 0;
  void setup(SetupContext& c) override {
    setupNodes(c);
    branch = c.allocBranch();
  }
  void load(LoadContext& lc) override {
    int a = u[0], b = u[1], k = branch;
    value = src.at(lc);
    lc.M->add(a, k, 1);
    lc.M->add(b, k, -1);
    lc.M->add(k, a, 1);
    lc.M->add(k, b, -1);
    lc.addRHS(k, value);
  }
  OperatingPoint operating(const double* x, int) override {
    OperatingPoint op;
    op.voltage = voltage(x);
    op.current = x[branch];
    op.power = op.voltage * op.current;
    return op;
  }
  bool breakpointer() const override { return true; }
  double nextBreakpoint(double t) const override { return src.nextBreakpoint(t); }
  void ports(const OperatingPoint& op, Accumulator& acc) override { acc.addPort(0, -op.power); }
  void loadAC(ACContext& ac) override {
    int a = u[0], b = u[1], k = branch;
    ac.M->add(a, k, 1);
    ac.M->add(b, k, -1);
    ac.M->add(k, a, 1);
    ac.M->add(k, b, -1);
    ac.addRHS(k, src.ac);
  }
};

// Controlled source, device/controlled.go:13-161: E (VCVS) and H (CCVS) force
// v(n+) - v(n-) through a branch unknown of their own, G (VCCS) and F (CCCS)
// drive a current n+ -> n-; E and G sense v(nc+) - v(nc-), F and H the branch
// current of another element (ctrl: its index, linked after setup). Linear,
// no state; all of its power is model-supplied.
struct Controlled : Device {
  double gain = 0;
  int cnode[2] = {0, 0};
  int ctrl = -1;
  int cu[2] = {-1, -1}, ck = -1, branch = -1;
  bool vout() const { return kind == KH_ELEM_E || kind == KH_ELEM_H; }
  bool vctl() const { return kind == KH_ELEM_E || kind == KH_ELEM_G; }
  void setup(SetupContext& c) override {
    setupNodes(c);
    if (vctl()) {
      cu[0] = c.unknown(cnode[0]);
      cu[1] = c.unknown(cnode[1]);
    }
    if (vout()) branch = c.allocBranch();
  }
  template <class M, class T>
  void stamp(M* m, T one, T g) {  // :80-114
    const int p = u[0], n = u[1];
    if (vout()) {
      m->add(p, branch, one);
      m->add(n, branch, -one);
      m->add(branch, p, one);
      m->add(branch, n, -one);
      if (kind == KH_ELEM_E) {
        m->add(branch, cu[0], -g);
        m->add(branch, cu[1], g);
      } else {
        m->add(branch, ck, -g);
      }
      return;
    }
    if (kind == KH_ELEM_G) {
      m->add(p, cu[0], g);
      m->add(p, cu[1], -g);
      m->add(n, cu[0], -g);
      m->add(n, cu[1], g);
    } else {
      m->add(p, ck, g);
      m->add(n, ck, -g);
    }
  }
  void load(LoadContext& lc) override { stamp(lc.M, 1.0, gain); }
  double control(const double* x) const {  // :116-121
    auto at = [&](int i) { return i < 0 ? 0.0 : x[i]; };
    return vctl() ? at(cu[0]) - at(cu[1]) : at(ck);
  }
  OperatingPoint operating(const double* x, int) override {  // :123-133
    OperatingPoint op;
    op.voltage = voltage(x);
    op.current = vout() ? x[branch] : gain * control(x);
    op.power = op.voltage * op.current;
    return op;
  }
  void ports(const OperatingPoint&, Accumulator&) override {}  // no exterior port of its own
  bool modelSupplier() const override { return true; }
  double modelSupplied(const OperatingPoint& op) const override { return -op.power; }
  void loadAC(ACContext& ac) override { stamp(ac.M, cplx(1, 0), cplx(gain, 0)); }
  cplx currentAC(const cplx* x) const {  // OperatingAC, :146-161
    auto at = [&](int i) { return i < 0 ? cplx(0) : x[i]; };
    if (vout()) return at(branch);
    const cplx ctl = kind == KH_ELEM_G ? at(cu[0]) - at(cu[1]) : at(ck);
    return cplx(gain, 0) * ctl;
  }
};

// CurrentSource, device/source.go:133-177, :207-217
struct CurrentSource : Device {
  SourceValue src;
  double value = 0;
  void load(LoadContext& lc) override {
    value = src.at(lc);
    lc.addRHS(u[0], -value);
    lc.addRHS(u[1], value);
  }
  OperatingPoint operating(const double* x, int) override {
    OperatingPoint op;
    op.voltage = voltage(x);
    op.current = value;
    op.power = op.voltage * value;
    return op;
  }
  bool breakpointer() const override { return true; }
  double nextBreakpoint(double t) const override { return src.nextBreakpoint(t); }
  void ports(const OperatingPoint& op, Accumulator& acc) override { acc.addPort(0, -op.power); }
  void loadAC(ACContext& ac) override {
    ac.addRHS(u[0], -src.ac);
    ac.addRHS(u[1], src.ac);
  }
};

// pnjlim, device/diode.go:214-226
inline double pnjlim(double vnew, double vold, double vt, double vcrit, bool* lim) {
  if (vnew > vcrit && std::fabs(vnew - vold) > 2 * vt) {
    *lim = true;
    if (vold > 0) {
      double arg = 1 + (vnew - vold) / vt;
      if (arg > 0) return vold + vt * std::log(arg);
      return vcrit;
    }
    return vt * std::log(vnew / vt);
  }
  *lim = false;
  return vnew;
}

// Diode, device/diode.go:61-435
struct Diode : Device {
  double IS = 0, N = 1, RS = 0, CJO = 0, VJ = 1, M = 0.5, TT = 0, BV = 0, IBV = 1e-3, FC = 0.5;
  bool hasBV = false;
  double vte = 0, vcrit = 0, xbv = 0, grs = 0;
  bool hasCharge = false;
  int ai = -1;
  double vd = 0, cd = 0, gd = 0, iq = 0, geq = 0;
  bool tran = false, lim = false;
  double gmin = 1e-12, qPrev = 0, iqPrev = 0, vdAcc = 0, omega = 0;
  History hist;

  void setup(SetupContext& c) override {
    setupNodes(c);
    ai = u[0];
    if (grs > 0) ai = c.allocUnknown();
  }
  void current(double v, double* id, double* g) const {  // :148-162
    if (v >= -3 * vte) {
      double e = std::exp(std::min(v / vte, 700.0));
      *id = IS * (e - 1);
      *g = IS * e / vte;
    } else if (!hasBV || v >= -xbv) {
      double a = 3 * vte / (v * kE);
      a = a * a * a;
      *id = -IS * (1 + a);
      *g = IS * 3 * a / v;
    } else {
      double e = std::exp(std::min(-(xbv + v) / vte, 700.0));
      *id = -IS * e;
      *g = IS * e / vte;
    }
  }
  void charge(double v, double id, double g, double* q, double* c) const {  // :191-211
    *q = TT * id;
    *c = TT * g;
    if (CJO == 0) return;
    double fcv = FC * VJ;
    if (v < fcv) {
      double arg = 1 - v / VJ;
      double sarg = std::pow(arg, -M);
      *q += CJO * VJ * (1 - arg * sarg) / (1 - M);
      *c += CJO * sarg;
      return;
    }
    double f1 = VJ * (1 - std::pow(1 - FC, 1 - M)) / (1 - M);
    double f2 = std::pow(1 - FC, 1 + M);
    double f3 = 1 - FC * (1 + M);
    *q += CJO * f1 + CJO / f2 * (f3 * (v - fcv) + M / (2 * VJ) * (v * v - fcv * fcv));
    *c += CJO / f2 * (f3 + M * v / VJ);
  }
  void load(LoadContext& lc) override {  // :229-275
    int a = u[0], k = u[1];
    gmin = lc.gmin;
    double vraw = at(lc.X, lc.nx, ai) - at(lc.X, lc.nx, k);
    if (lc.initJct) {
      vd = vcrit;
      lim = true;
    } else {
      vd = pnjlim(vraw, vdAcc, vte, vcrit, &lim);
      if (hasBV && vraw < std::min(0.0, -xbv + 10 * vte)) {
        bool l2;
        double vt = pnjlim(-(vraw + xbv), -(vdAcc + xbv), vte, vcrit, &l2);
        vd = -(vt + xbv);
        lim = l2;
      }
    }
    vdAcc = vd;
    double id, g0;
    current(vd, &id, &g0);
    cd = id + gmin * vd;
    gd = g0 + gmin;
    tran = false; iq = 0; geq = 0;
    if (lc.mode == ModeTran && hasCharge) {
      double q, c;
      charge(vd, id, g0, &q, &c);
      tran = true;
      if (lc.method == Trapezoidal) {
        iq = 2 * (q - qPrev) / lc.H - iqPrev;
        geq = 2 * c / lc.H;
      } else {
        iq = (q - qPrev) / lc.H;
        geq = c / lc.H;
      }
    }
    double g = gd + geq;
    double ieq = cd + iq - g * vd;
    lc.M->add(ai, ai, g);
    lc.M->add(k, k, g);
    lc.M->add(ai, k, -g);
    lc.M->add(k, ai, -g);
    lc.addRHS(ai, -ieq);
    lc.addRHS(k, ieq);
    if (grs > 0) {
      lc.M->add(a, a, grs);
      lc.M->add(ai, ai, grs);
      lc.M->add(a, ai, -grs);
      lc.M->add(ai, a, -grs);
    }
  }
  bool nonlinear() const override { return true; }
  bool limited() const override { return lim; }
  bool convTest(const double* x, int nx, double reltol, double abstol) override {  // :282-289
    double v = at(x, nx, ai) - at(x, nx, u[1]);
    double cdhat = cd + gd * (v - vd);
    double id, g0;
    current(v, &id, &g0);
    double cdn = id + gmin * v;
    double tol = reltol * std::max(std::fabs(cdhat), std::fabs(cdn)) + abstol;
    return std::fabs(cdhat - cdn) <= tol;
  }
  OperatingPoint operating(const double* x, int nx) override {  // :293-301
    double v = at(x, nx, ai) - at(x, nx, u[1]);
    double ir = cd + gd * (v - vd);
    double ic = iq + geq * (v - vd);
    OperatingPoint op;
    op.current = ir + ic;
    op.voltage = voltage(x);
    op.power = op.voltage * op.current;
    op.storage = v * ic;
    return op;
  }
  void accept(const double* x, int nx, double t, int mode) override {  // :304-322
    double v = at(x, nx, ai) - at(x, nx, u[1]);
    vdAcc = v;
    double id, g0, q, c;
    current(v, &id, &g0);
    charge(v, id, g0, &q, &c);
    if (mode == ModeTran) {
      iqPrev = tran ? iq + geq * (v - vd) : 0;
    } else {
      iqPrev = 0;
      hist.reset();
    }
    qPrev = q;
    hist.push(t, q);
  }
  bool reactive() const override { return true; }
  double truncate(const double* x, int nx, const TruncContext& tc) override {  // :388-397
    if (!hasCharge) return kInf;
    double v = at(x, nx, ai) - at(x, nx, u[1]);
    double id, g0, q, c;
    current(v, &id, &g0);
    charge(v, id, g0, &q, &c);
    double ic = iq + geq * (v - vd);
    return lteStep(hist, q, ic, iqPrev, tc.absTol, tc);
  }
  // StoredEnergy, device/diode.go:325-383: vd·q(vd) − ∫₀^vd qj − TT·∫₀^vd id (closed forms)
  double intDepletion(double v) const {
    if (CJO == 0) return 0;
    double fcv = FC * VJ;
    auto below = [&](double w) { return CJO * VJ / (1 - M) * (w + VJ / (2 - M) * (std::pow(1 - w / VJ, 2 - M) - 1)); };
    if (v < fcv) return below(v);
    double f1 = VJ * (1 - std::pow(1 - FC, 1 - M)) / (1 - M);
    double f2 = std::pow(1 - FC, 1 + M);
    double f3 = 1 - FC * (1 + M);
    double dv = v - fcv;
    return below(fcv) + CJO * f1 * dv + CJO / f2 * (f3 * dv * dv / 2 + M / (2 * VJ) * ((v * v * v - fcv * fcv * fcv) / 3 - fcv * fcv * dv));
  }
  double intCurrent(double v) const {
    auto fwd = [&](double w) { return IS * (vte * std::expm1(std::min(w / vte, 700.0)) - w); };
    double v1 = -3 * vte;
    if (v >= v1) return fwd(v);
    double total = fwd(v1);
    double a3 = std::pow(3 * vte / kE, 3.0);
    auto rev = [&](double w) { return -IS * (w - a3 / (2 * w * w)); };
    double v2 = v;
    if (hasBV && v < -xbv) v2 = -xbv;
    total += rev(v2) - rev(v1);
    if (v2 == v) return total;
    auto brk = [&](double w) { return IS * vte * std::exp(std::min(-(xbv + w) / vte, 700.0)); };
    return total + brk(v) - brk(v2);
  }
  double storedEnergy(const double* x, int nx) override {
    if (!hasCharge) return 0;
    double v = at(x, nx, ai) - at(x, nx, u[1]);
    double id, g0, q, c;
    current(v, &id, &g0);
    charge(v, id, g0, &q, &c);
    return v * q - intDepletion(v) - TT * intCurrent(v);
  }
  void ports(const OperatingPoint& op, Accumulator& acc) override { acc.addPort(2, -(op.power - op.storage)); }
  void loadAC(ACContext& ac) override {  // :409-420
    int a = u[0], k = u[1];
    double id, g0, q, c;
    current(vdAcc, &id, &g0);
    charge(vdAcc, id, g0, &q, &c);
    stampY(ac.M, ai, k, cplx(g0 + gmin, ac.omega * c));
    if (grs > 0) stampY(ac.M, a, ai, cplx(grs, 0));
    omega = ac.omega;
  }
};

// fetlim, device/mosfet.go:347-386
inline double fetlim(double vnew, double vold, double vto) {
  double vtsthi = std::fabs(2 * (vold - vto)) + 2;
  double vtstlo = vtsthi / 2 + 2;
  double vtox = vto + 3.5;
  double delv = vnew - vold;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) {
          if (-delv > vtstlo) vnew = vold - vtstlo;
        } else {
          vnew = std::max(vnew, vto + 2);
        }
      } else if (delv >= vtsthi) {
        vnew = vold + vtsthi;
      }
    } else if (delv <= 0) {
      vnew = std::max(vnew, vto - 0.5);
    } else {
      vnew = std::min(vnew, vto + 4);
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

// limvds, device/mosfet.go:388-404
inline double limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) return std::min(vnew, 3 * vold + 2);
    if (vnew < 3.5) return std::max(vnew, 2.0);
    return vnew;
  }
  if (vnew > vold) return std::min(vnew, 4.0);
  return std::max(vnew, -0.5);
}

// depletion, device/bjt.go:358-370
inline void depletion(double v, double cz, double vj, double m, double fc, double f1, double f2, double f3,
                      double* qv, double* c) {
  if (cz == 0) { *qv = 0; *c = 0; return; }
  double fcv = fc * vj;
  if (v < fcv) {
    double arg = 1 - v / vj;
    double sarg = std::exp(-m * std::log(arg));
    *qv = vj * cz * (1 - arg * sarg) / (1 - m);
    *c = cz * sarg;
    return;
  }
  double czf2 = cz / f2;
  *qv = cz * f1 + czf2 * (f3 * (v - fcv) + (m / (vj + vj)) * (v * v - fcv * fcv));
  *c = czf2 * (f3 + m * v / vj);
}

// meyer, device/mosfet.go:419-438
inline void meyer(double vgs, double vgd, double von, double vdsat, double phi, double cox, double* cgs, double* cgd,
                  double* cgb) {
  double vgst = vgs - von;
  vdsat = std::max(vdsat, 0.0);
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

constexpr double EPS_SI = 1.035943139907e-10;  // epsSi = 11.7 * 8.854214871e-12, folded exactly

// Forward-mode dual number over (vgs, vds, vbs), device/dual.go:17-106: the
// MOSFET level 2 and 3 currents are evaluated in it, so gm, gds and gmbs are
// their exact derivatives. Each operation restates the Go method's arithmetic.
struct Dual {
  double v = 0, d[3] = {0, 0, 0};
};
inline Dual dcst(double v) { Dual r; r.v = v; return r; }
inline Dual dvar(double v, int k) { Dual r; r.v = v; r.d[k] = 1; return r; }
inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r; r.v = a.v + b.v;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r; r.v = a.v - b.v;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
inline Dual addc(Dual a, double c) { a.v += c; return a; }
inline Dual dneg(const Dual& a) {
  Dual r; r.v = -a.v;
  for (int k = 0; k < 3; k++) r.d[k] = -a.d[k];
  return r;
}
inline Dual scale(const Dual& a, double c) {
  Dual r; r.v = a.v * c;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] * c;
  return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r; r.v = a.v * b.v;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
  Dual r; double q = a.v / b.v; r.v = q;
  for (int k = 0; k < 3; k++) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
  return r;
}
inline Dual chain(const Dual& a, double f, double df) {
  Dual r; r.v = f;
  for (int k = 0; k < 3; k++) r.d[k] = df * a.d[k];
  return r;
}
inline Dual dsqrt(const Dual& a) { double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
inline Dual dexp(const Dual& a) { double e = std::exp(a.v); return chain(a, e, e); }
inline Dual dlog(const Dual& a) { return chain(a, std::log(a.v), 1 / a.v); }
inline Dual dcos(const Dual& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline Dual datan(const Dual& a) { return chain(a, std::atan(a.v), 1 / (1 + a.v * a.v)); }
inline Dual dabs(const Dual& a) { return a.v < 0 ? dneg(a) : a; }
inline Dual powc(const Dual& a, double c) { double p = std::pow(a.v, c); return chain(a, p, c * p / a.v); }
inline Dual dmax(const Dual& a, const Dual& b) { return b.v > a.v ? b : a; }
inline Dual dmin(const Dual& a, const Dual& b) { return b.v < a.v ? b : a; }

// MOSFET level 2 and 3 parameters and derived constants (the M layout's
// level block, device/mosfet.go:128-324 and device/mosfet_levels.go:22-55).
struct MosLevel {
  int level = 1;
  double NFS = 0, XJ = 0, UCRIT = 1e4, UEXP = 0, VMAX = 0, NEFF = 1, DELTA = 0, THETA = 0, ETA = 0, KAPPA = 0, UO = 600;
  double leff = 0, weff = 0, cox = 0, xd = 0, alpha = 0, cfs = 0, factor = 0, narrow = 0, eta3 = 0, sqrtPB = 0;
  bool hasNSUB = false;
};

struct IdsOut {
  double cd = 0, gm = 0, gds = 0, gmbs = 0, von = 0, vdsat = 0;
};
inline IdsOut fromDual(const Dual& id, const Dual& von, const Dual& vdsat) {
  IdsOut o;
  o.cd = id.v; o.gm = id.d[0]; o.gds = id.d[1]; o.gmbs = id.d[2]; o.von = von.v; o.vdsat = vdsat.v;
  return o;
}

// cbrtAbs, device/mosfet_level23.go:263-270
inline Dual cbrtAbs(Dual a) {
  a = dabs(a);
  if (a.v == 0) return Dual();
  return powc(a, 1.0 / 3);
}

// level2VelocitySat, device/mosfet_level23.go:195-261: the Baum–Beneking
// quartic for x = sqrt(vdsat + phi - vbs), Ferrari's closed form; the smallest
// positive root with |poly| <= 1e-6.
inline bool level2VelocitySat(const Dual& gammad, const Dual& v1, const Dual& v2, const Dual& invxv, const Dual& sarg3,
                              Dual* xout) {
  Dual xv = dcst(1) / invxv;
  Dual a1 = scale(gammad, 1 / 0.75);
  Dual b1 = scale(v1 + xv, -2);
  Dual c1 = scale(gammad * xv, -2);
  Dual d1 = scale(v1 * (v2 + xv), 2) - v2 * v2 - scale(gammad * sarg3, 4.0 / 3.0);
  Dual a = dneg(b1);
  Dual b = a1 * c1 - scale(d1, 4);
  Dual c = dneg(d1) * (a1 * a1 - scale(b1, 4)) - c1 * c1;
  Dual r = scale(a * a, -1.0 / 3) + b;
  Dual s = scale(a * a * a, 2.0 / 27) - scale(a * b, 1.0 / 3) + c;
  Dual r3 = r * r * r;
  Dual s2 = s * s;
  Dual pp = scale(s2, 0.25) + scale(r3, 1.0 / 27);
  Dual p0 = dabs(pp);
  Dual p2 = dsqrt(p0);
  Dual y3;
  if (pp.v < 0) {
    Dual ro = dexp(scale(dlog(dsqrt(scale(s2, 0.25) + p0)), 1.0 / 3));
    Dual fi = datan(scale(p2, -2) / s);
    y3 = scale(ro * dcos(scale(fi, 1.0 / 3)), 2) - scale(a, 1.0 / 3);
  } else {
    y3 = cbrtAbs(scale(s, -0.5) + p2) + cbrtAbs(scale(s, -0.5) - p2) - scale(a, 1.0 / 3);
  }
  Dual ra = scale(a1 * a1, 0.25) - b1 + y3;
  Dual rb = scale(y3 * y3, 0.25) - d1;
  if (!(ra.v >= 0) || !(rb.v >= 0)) return false;
  Dual a3 = dsqrt(ra), b3 = dsqrt(rb);
  const double sig1[4] = {1, -1, 1, -1}, sig2[4] = {1, 1, -1, -1};
  bool ok = false;
  Dual x;
  for (int i = 0; i < 4; i++) {
    Dual a4 = scale(a1, 0.5) + scale(a3, sig1[i]);
    Dual b4 = scale(y3, 0.5) + scale(b3, sig2[i]);
    Dual delta4 = scale(a4 * a4, 0.25) - b4;
    if (!(delta4.v >= 0)) continue;
    Dual t = dsqrt(delta4);
    Dual roots[2] = {scale(a4, -0.5) + t, scale(a4, -0.5) - t};
    for (const Dual& rt : roots) {
      if (!(rt.v > 0)) continue;
      double x2 = rt.v * rt.v;
      double poly = x2 * x2 + a1.v * x2 * rt.v + b1.v * x2 + c1.v * rt.v + d1.v;
      if (!(std::fabs(poly) <= 1e-6)) continue;
      if (!ok || rt.v < x.v) { x = rt; ok = true; }
    }
  }
  *xout = x;
  return ok;
}

// idsLevel2, device/mosfet_level23.go:31-193 (SPICE2 level 2, Grove–Frohman).
inline IdsOut idsLevel2(const MosLevel& lv, double vbi, double beta, double vt, double PHI, double GAMMA,
                        double LAMBDA, double vgsV, double vdsV, double vbsV) {
  Dual G = dvar(vgsV, 0), D = dvar(vdsV, 1), B = dvar(vbsV, 2);
  double phi = PHI;
  double sphi = std::sqrt(phi);
  double sphi3 = phi * sphi;
  Dual phiMinVbs = addc(dneg(B), phi);
  Dual sarg, dsrgdb;
  if (B.v <= 0) {
    sarg = dsqrt(phiMinVbs);
    dsrgdb = dcst(-0.5) / sarg;
  } else {
    sarg = dcst(sphi) / addc(scale(B, 0.5 / phi), 1);
    dsrgdb = scale(sarg * sarg, -0.5 / sphi3);
  }
  Dual barg, dbrgdb;
  if (B.v - D.v <= 0) {
    barg = dsqrt(phiMinVbs + D);
    dbrgdb = dcst(-0.5) / barg;
  } else {
    barg = dcst(sphi) / addc(scale(B - D, 0.5 / phi), 1);
    dbrgdb = scale(barg * barg, -0.5 / sphi3);
  }
  double factor = lv.factor;
  double eta = 1 + factor;
  Dual vbin = addc(scale(phiMinVbs, factor), vbi);
  Dual gamasd = dcst(GAMMA), dgddvb = dcst(0);
  if (GAMMA > 0 || lv.hasNSUB) {
    if (lv.XJ > 0) {
      Dual xws = scale(sarg, lv.xd), xwd = scale(barg, lv.xd);
      Dual args = dsqrt(addc(scale(xws, 2 / lv.XJ), 1));
      Dual argd = dsqrt(addc(scale(xwd, 2 / lv.XJ), 1));
      double t = 0.5 * lv.XJ / lv.leff;
      Dual argss = scale(addc(args, -1), t), argsd = scale(addc(argd, -1), t);
      gamasd = scale(addc(dneg(argss + argsd), 1), GAMMA);
      Dual dbargs = scale(scale(dsrgdb, lv.xd) / args, 0.5 / lv.leff);
      Dual dbargd = scale(scale(dbrgdb, lv.xd) / argd, 0.5 / lv.leff);
      dgddvb = scale(dbargs + dbargd, -GAMMA);
    }
  }
  Dual vonD = vbin + gamasd * sarg;
  bool nfs = lv.NFS != 0;
  Dual xn, argg;
  if (nfs) {
    Dual cdonco = addc(dneg(gamasd * dsrgdb + dgddvb * sarg), factor);
    xn = addc(cdonco, 1 + lv.cfs);
    Dual vtxn = scale(xn, vt);
    vonD = vonD + vtxn;
    argg = dcst(1) / vtxn;
  } else if (G.v <= vbin.v) {
    IdsOut o;
    o.von = vonD.v;
    return o;  // cutoff
  }
  Dual vgst = G - vonD;
  Dual ufact = dcst(1);
  double crit = lv.UCRIT * 100 * EPS_SI / lv.cox;
  if (vgst.v > crit) ufact = powc(dcst(crit) / vgst, lv.UEXP);
  Dual ueff = scale(ufact, lv.UO * 1e-4);
  Dual vgsx = G;
  if (nfs) vgsx = dmax(G, vonD);
  Dual gammad = scale(gamasd, 1 / eta);
  Dual vgb = scale(vgsx - vbin, 1 / eta);
  Dual vdsatD = dcst(0);
  if (gammad.v > 0) {
    Dual argv = vgb + phiMinVbs;
    if (argv.v > 0) {
      Dual gammd2 = gammad * gammad;
      Dual arg1 = dsqrt(addc(scale(argv, 4) / gammd2, 1));
      vdsatD = dmax(vgb + scale(gammd2 * addc(dneg(arg1), 1), 0.5), dcst(0));
    }
  } else {
    vdsatD = dmax(vgb, dcst(0));
  }
  Dual sarg3 = sarg * sarg * sarg;
  if (lv.VMAX > 0) {
    Dual x;
    if (level2VelocitySat(gammad, vgb + phiMinVbs, phiMinVbs, scale(ueff, 1 / (lv.VMAX * lv.leff)), sarg3, &x))
      vdsatD = x * x - phiMinVbs;
  }
  Dual xlamda = dcst(LAMBDA);
  Dual bsarg, bodys;
  if (D.v != 0) {
    if (B.v - vdsatD.v <= 0) bsarg = dsqrt(vdsatD + phiMinVbs);
    else bsarg = dcst(sphi) / addc(scale(B - vdsatD, 0.5 / phi), 1);
    bodys = bsarg * bsarg * bsarg - sarg3;
    if (lv.hasNSUB && !(LAMBDA > 0)) {
      if (lv.VMAX <= 0) {
        Dual argv = scale(D - vdsatD, 0.25);
        Dual sargv = dsqrt(addc(argv * argv, 1));
        xlamda = dsqrt(argv + sargv) * (dcst(lv.xd) / scale(D, lv.leff));
      } else {
        double xdv = lv.xd / std::sqrt(lv.NEFF);
        Dual xlv = dcst(lv.VMAX * xdv * 0.5) / ueff;
        Dual argv = dmax(D - vdsatD, dcst(0));
        Dual xls = dsqrt(xlv * xlv + argv);
        xlamda = (xls - xlv) * (dcst(xdv) / scale(D, lv.leff));
      }
    }
  }
  double xwb = lv.xd * lv.sqrtPB;
  double xld = lv.leff - xwb;
  Dual clfact = addc(dneg(xlamda * D), 1);
  if (!lv.hasNSUB) xwb = 0.25e-6;
  Dual xleff = scale(clfact, lv.leff);
  if (xleff.v < xwb) {
    Dual deltal = scale(xlamda * D, lv.leff);
    xleff = dcst(xwb) / addc(scale(addc(deltal, -xld), 1 / xwb), 1);
    clfact = scale(xleff, 1 / lv.leff);
  }
  Dual beta1 = scale(ufact, beta) / clfact;
  if (D.v <= 1e-10) {
    Dual g;
    if (G.v > vonD.v) g = beta1 * (G - vbin - gamasd * sarg);
    else if (nfs) g = beta1 * (vonD - vbin - gamasd * sarg) * dexp(argg * (G - vonD));
    IdsOut o;
    o.gds = g.v; o.von = vonD.v; o.vdsat = vdsatD.v;
    return o;
  }
  Dual id;
  if (nfs && G.v <= vonD.v) {
    if (vdsatD.v <= 0) {
      IdsOut o;
      o.von = vonD.v; o.vdsat = vdsatD.v;
      return o;
    }
    Dual vdson = dmin(vdsatD, D);
    Dual body = barg * barg * barg - sarg3;
    if (D.v > vdsatD.v) body = bodys;
    Dual cdson = beta1 * ((vonD - vbin - scale(vdson, 0.5 * eta)) * vdson - scale(gamasd * body, 1 / 1.5));
    id = cdson * dexp(argg * (G - vonD));
  } else if (D.v <= vdsatD.v) {
    Dual body = barg * barg * barg - sarg3;
    id = beta1 * ((G - vbin - scale(D, 0.5 * eta)) * D - scale(gamasd * body, 1 / 1.5));
  } else {
    id = beta1 * ((G - vbin - scale(vdsatD, 0.5 * eta)) * vdsatD - scale(gamasd * bodys, 1 / 1.5));
  }
  return fromDual(id, vonD, vdsatD);
}

// idsLevel3, device/mosfet_level23.go:282-378 (SPICE2 level 3, semi-empirical).
inline IdsOut idsLevel3(const MosLevel& lv, double vbi, double beta, double vt, double PHI, double GAMMA,
                        double LD, double vgsV, double vdsV, double vbsV) {
  const double coeff0 = 0.0631353, coeff1 = 0.8013292, coeff2 = -0.01110777;
  Dual G = dvar(vgsV, 0), D = dvar(vdsV, 1), B = dvar(vbsV, 2);
  double phi = PHI;
  double leff = lv.leff, weff = lv.weff;
  Dual phibs, sqphbs;
  if (B.v <= 0) {
    phibs = addc(dneg(B), phi);
    sqphbs = dsqrt(phibs);
  } else {
    sqphbs = dcst(std::sqrt(phi)) / addc(scale(B, 1 / (2 * phi)), 1);
    phibs = sqphbs * sqphbs;
  }
  Dual fshort = dcst(1);
  if (lv.XJ != 0 && lv.xd != 0) {
    Dual wponxj = scale(sqphbs, lv.xd / lv.XJ);
    Dual wconxj = addc(scale(wponxj * wponxj, coeff2) + scale(wponxj, coeff1), coeff0);
    double djonxj = LD / lv.XJ;
    Dual arga = addc(wconxj, djonxj);
    Dual argc = wponxj / addc(wponxj, 1);
    Dual argb = dsqrt(addc(dneg(argc * argc), 1));
    fshort = addc(scale(addc(arga * argb, -djonxj), -lv.XJ / leff), 1);
  }
  Dual gammas = scale(fshort, GAMMA);
  Dual fbodys = scale(gammas / sqphbs, 0.25);
  Dual fbody = addc(fbodys, lv.narrow / weff);
  Dual qbonco = gammas * sqphbs + scale(phibs, lv.narrow / weff);
  Dual vth = addc(scale(D, -lv.eta3), vbi) + qbonco;
  Dual vonD = vth;
  bool nfs = lv.NFS != 0;
  Dual xn;
  if (nfs) {
    xn = addc(qbonco / scale(phibs, 2), 1 + lv.cfs);
    vonD = vth + scale(xn, vt);
  } else if (G.v <= vonD.v) {
    IdsOut o;
    o.von = vonD.v;
    return o;
  }
  Dual vgsx = dmax(G, vonD);
  Dual onfg = addc(scale(vgsx - vth, lv.THETA), 1);
  Dual fgate = dcst(1) / onfg;
  Dual us = scale(fgate, lv.UO * 1e-4);
  Dual onfbdy = dcst(1) / addc(fbody, 1);
  Dual arga = (vgsx - vth) * onfbdy;
  Dual vdsatD = arga;
  Dual vdsc;
  if (lv.VMAX > 0) {
    vdsc = dcst(leff * lv.VMAX) / us;
    Dual argb = dsqrt(arga * arga + vdsc * vdsc);
    vdsatD = arga + vdsc - argb;
  }
  Dual vdsx = dmin(D, vdsatD);
  Dual betaD = scale(fgate, beta);
  if (vdsx.v == 0) {
    Dual g = betaD * (vgsx - vth);
    if (nfs && G.v < vonD.v) g = g * dexp((G - vonD) / scale(xn, vt));
    IdsOut o;
    o.gds = g.v; o.von = vonD.v; o.vdsat = vdsatD.v;
    return o;
  }
  Dual cdo = vgsx - vth - scale(addc(fbody, 1) * vdsx, 0.5);
  Dual id = betaD * cdo * vdsx;
  Dual fdrain;
  if (lv.VMAX > 0) {
    fdrain = dcst(1) / addc(vdsx / vdsc, 1);
    id = id * fdrain;
  }
  bool clm = false;
  Dual delxl;
  if (D.v <= vdsatD.v) {
    if (lv.VMAX <= 0 && lv.alpha != 0) {
      Dual r = D / vdsatD;
      Dual r2 = r * r;
      delxl = dsqrt(scale(vdsatD, lv.KAPPA * lv.alpha / 8)) * (r2 * r2);
      clm = true;
    }
  } else if (lv.VMAX <= 0) {
    delxl = dsqrt(scale(D - vdsatD + scale(vdsatD, 1.0 / 8), lv.KAPPA * lv.alpha));
    clm = true;
  } else if (lv.alpha != 0) {
    Dual gdsat = dmax(id * addc(dneg(fdrain), 1) / vdsc, dcst(1e-12));
    Dual emax = scale(id / gdsat, lv.KAPPA / leff);
    Dual a = scale(emax, 0.5 * lv.alpha);
    Dual argb = dsqrt(a * a + scale(D - vdsatD, lv.KAPPA * lv.alpha));
    delxl = argb - a;
    clm = true;
  }
  if (clm) {
    if (delxl.v > 0.5 * leff) delxl = addc(dcst(-leff * leff / 4) / delxl, leff);
    id = id / addc(scale(delxl, -1 / leff), 1);
  }
  if (nfs && G.v < vonD.v) id = id * dexp((G - vonD) / scale(xn, vt));
  return fromDual(id, vonD, vdsatD);
}

// MOSFET levels 1-3, device/mosfet.go:43-846, device/mosfet_levels.go:79-113 and
// device/mosfet_level23.go
struct Mosfet : Device {
  double typ = 1, beta = 0, vbi = 0, vto = 0, GAMMA = 0, PHI = 0.6, LAMBDA = 0, gdpr = 0, gspr = 0;
  double isd = 0, iss = 0, vcritD = 0, vcritS = 0, czbd = 0, czbdsw = 0, czbs = 0, czbssw = 0;
  double PB = 0.8, MJ = 0.5, MJSW = 0.5, FC = 0.5, fbt[3] = {0, 0, 0}, fsw[3] = {0, 0, 0};
  double cgso = 0, cgdo = 0, cgbo = 0, oxideCap = 0, vt = 0;
  bool hasJcap = false, hasGcap = false;
  MosLevel lv;
  double LD = 0;
  int g = -1, b = -1, dp = -1, sp = -1;
  double vgs = 0, vds = 0, vbs = 0, vbd = 0, vgd = 0, vgb = 0;
  double mode = 1, cdrain = 0, gm = 0, gds = 0, gmbs = 0, von = 0, vdsat = 0;
  double cbs = 0, gbs = 0, cbd = 0, gbd = 0;
  bool lim = false, tran = false;
  double ag = 0, gmin = 1e-12;
  double capG[3] = {0, 0, 0}, capB[2] = {0, 0}, iq[5] = {0, 0, 0, 0, 0}, dv[5] = {0, 0, 0, 0, 0};
  struct { double vgs = 0, vds = 0, vbs = 0, vgd = 0, vgb = 0, vbd = 0, von = 0; } acc;
  double meyerAcc[3] = {0, 0, 0}, qPrev[5] = {0, 0, 0, 0, 0}, iqPrev[5] = {0, 0, 0, 0, 0}, vPrev[5] = {0, 0, 0, 0, 0};
  History hist[5];
  struct { double gch[3] = {0, 0, 0}, gbs = 0, gbd = 0, cap[5] = {0, 0, 0, 0, 0}, omega = 0; } ac;

  void setup(SetupContext& c) override {  // :326-337
    setupNodes(c);
    g = c.unknown(nodes[2]);
    b = c.unknown(nodes[3]);
    dp = u[0];
    sp = u[1];
    if (gdpr > 0) dp = c.allocUnknown();
    if (gspr > 0) sp = c.allocUnknown();
  }
  void voltages(const double* x, int nx, double* ovgs, double* ovds, double* ovbs, double* ovgb) const {
    double t = typ;
    double vg = at(x, nx, g), vb = at(x, nx, b), vdp = at(x, nx, dp), vsp = at(x, nx, sp);
    *ovgs = t * (vg - vsp);
    *ovds = t * (vdp - vsp);
    *ovbs = t * (vb - vsp);
    *ovgb = t * (vg - vb);
  }
  void bulkDiode(double v, double is, double* i, double* gg) const {  // :408-415
    if (v <= -3 * vt) { *i = gmin * v - is; *gg = gmin; return; }
    double e = std::exp(std::min(v / vt, 700.0));
    *i = is * (e - 1) + gmin * v;
    *gg = is * e / vt + gmin;
  }
  void meyerCaps(double a_vgs, double a_vgd, double a_von, double a_vdsat, double md, double out[3]) const {  // :441-451
    if (oxideCap == 0) { out[0] = out[1] = out[2] = 0; return; }
    double a, bb, c;
    if (md > 0) {
      meyer(a_vgs, a_vgd, a_von, a_vdsat, PHI, oxideCap, &a, &bb, &c);
      out[0] = a; out[1] = bb; out[2] = c;
      return;
    }
    meyer(a_vgd, a_vgs, a_von, a_vdsat, PHI, oxideCap, &a, &bb, &c);
    out[0] = bb; out[1] = a; out[2] = c;
  }
  void bulkCharge(double v, double cz, double czsw, double* qv, double* c) const {  // :455-460
    double q1, c1, q2, c2;
    depletion(v, cz, PB, MJ, FC, fbt[0], fbt[1], fbt[2], &q1, &c1);
    depletion(v, czsw, PB, MJSW, FC, fsw[0], fsw[1], fsw[2], &q2, &c2);
    *qv = q1 + q2;
    *c = c1 + c2;
  }
  // ids, device/mosfet_levels.go:58-68: the level's drain current
  void ids(double a_vgs, double a_vds, double a_vbs, double* cd, double* ogm, double* ogds, double* ogmbs, double* ovon,
           double* ovdsat) const {
    if (lv.level == 2 || lv.level == 3) {
      IdsOut o = lv.level == 2 ? idsLevel2(lv, vbi, beta, vt, PHI, GAMMA, LAMBDA, a_vgs, a_vds, a_vbs)
                               : idsLevel3(lv, vbi, beta, vt, PHI, GAMMA, LD, a_vgs, a_vds, a_vbs);
      *cd = o.cd; *ogm = o.gm; *ogds = o.gds; *ogmbs = o.gmbs; *ovon = o.von; *ovdsat = o.vdsat;
      return;
    }
    // idsLevel1, device/mosfet_levels.go:79-113
    double sarg;
    if (a_vbs <= 0) {
      sarg = std::sqrt(PHI - a_vbs);
    } else {
      sarg = std::sqrt(PHI);
      sarg = std::max(0.0, sarg - a_vbs / (sarg + sarg));
    }
    double v_on = vbi + GAMMA * sarg;
    double vgst = a_vgs - v_on;
    *ovon = v_on;
    *ovdsat = std::max(vgst, 0.0);
    double arg = 0;
    if (sarg > 0) arg = GAMMA / (sarg + sarg);
    if (vgst <= 0) { *cd = 0; *ogm = 0; *ogds = 0; *ogmbs = 0; return; }
    double betap = beta * (1 + LAMBDA * a_vds);
    if (vgst <= a_vds) {
      *cd = betap * vgst * vgst * 0.5;
      *ogm = betap * vgst;
      *ogds = LAMBDA * beta * vgst * vgst * 0.5;
    } else {
      *cd = betap * a_vds * (vgst - 0.5 * a_vds);
      *ogm = betap * a_vds;
      *ogds = betap * (vgst - a_vds) + LAMBDA * beta * a_vds * (vgst - 0.5 * a_vds);
    }
    *ogmbs = *ogm * arg;
  }
  // channel, device/mosfet.go:464-473
  void channel(double a_vgs, double a_vds, double a_vbs, double* md, double* cd, double* ogm, double* ogds, double* ogmbs,
               double* ovon, double* ovdsat) const {
    *md = 1;
    double vg = a_vgs, vd = a_vds, vb = a_vbs;
    if (a_vds < 0) {
      *md = -1;
      vg = a_vgs - a_vds;
      vd = -a_vds;
      vb = a_vbs - a_vds;
    }
    ids(vg, vd, vb, cd, ogm, ogds, ogmbs, ovon, ovdsat);
  }
  double capk(int k) const { return k < 3 ? capG[k] : capB[k - 3]; }
  // branch, device/mosfet.go:563-578
  void branch(LoadContext& lc, int a, int bb, double i0, int n, const double gg[3], const int p[3], const int m[3],
              const double v0[3]) {
    double ieq = i0;
    for (int k = 0; k < n; k++) {
      if (gg[k] == 0) continue;
      lc.M->add(a, p[k], gg[k]);
      lc.M->add(a, m[k], -gg[k]);
      lc.M->add(bb, p[k], -gg[k]);
      lc.M->add(bb, m[k], gg[k]);
      ieq -= gg[k] * v0[k];
    }
    ieq *= typ;
    lc.addRHS(a, -ieq);
    lc.addRHS(bb, ieq);
  }
  void channelBranch(double* i0, double gc[3]) const {  // :582-589
    if (mode > 0) { *i0 = cdrain; gc[0] = gm; gc[1] = gds; gc[2] = gmbs; return; }
    *i0 = -cdrain;
    gc[0] = -gm;
    gc[1] = gm + gds + gmbs;
    gc[2] = -gmbs;
  }
  static void stampG(Stamper* m, int a, int bb, double gg) {  // device/bjt.go:482-490
    if (gg == 0) return;
    m->add(a, a, gg);
    m->add(bb, bb, gg);
    m->add(a, bb, -gg);
    m->add(bb, a, -gg);
  }
  void stamp(LoadContext& lc) {  // :591-611
    double i0, gc[3];
    channelBranch(&i0, gc);
    { int p[3] = {g, dp, b}, m[3] = {sp, sp, sp}; double v0[3] = {vgs, vds, vbs}; branch(lc, dp, sp, i0, 3, gc, p, m, v0); }
    { double gg[3] = {gbs, 0, 0}; int p[3] = {b, 0, 0}, m[3] = {sp, 0, 0}; double v0[3] = {vbs, 0, 0}; branch(lc, b, sp, cbs, 1, gg, p, m, v0); }
    { double gg[3] = {gbd, 0, 0}; int p[3] = {b, 0, 0}, m[3] = {dp, 0, 0}; double v0[3] = {vbd, 0, 0}; branch(lc, b, dp, cbd, 1, gg, p, m, v0); }
    if (tran) {
      int ends[5][2] = {{g, sp}, {g, dp}, {g, b}, {b, sp}, {b, dp}};
      for (int k = 0; k < 5; k++) {
        double c = ag * capk(k);
        double gg[3] = {c, 0, 0};
        int p[3] = {ends[k][0], 0, 0}, m[3] = {ends[k][1], 0, 0};
        double v0[3] = {dv[k], 0, 0};
        branch(lc, ends[k][0], ends[k][1], iq[k], 1, gg, p, m, v0);
      }
    }
    if (dp != u[0]) stampG(lc.M, u[0], dp, gdpr);
    if (sp != u[1]) stampG(lc.M, u[1], sp, gspr);
  }
  void load(LoadContext& lc) override {  // :476-551
    gmin = lc.gmin;
    double a_vgs, a_vds, a_vbs, a_vgb;
    voltages(lc.X, lc.nx, &a_vgs, &a_vds, &a_vbs, &a_vgb);
    lim = false;
    if (lc.initJct) {
      a_vbs = -1; a_vgs = vto; a_vds = 0;
      lim = true;
    } else {
      double a_vgd = a_vgs - a_vds;
      double vgdOld = acc.vgs - acc.vds;
      double v_on = acc.von;
      if (acc.vds >= 0) {
        a_vgs = fetlim(a_vgs, acc.vgs, v_on);
        a_vds = a_vgs - a_vgd;
        a_vds = limvds(a_vds, acc.vds);
        a_vgd = a_vgs - a_vds;
      } else {
        a_vgd = fetlim(a_vgd, vgdOld, v_on);
        a_vds = a_vgs - a_vgd;
        a_vds = -limvds(-a_vds, -acc.vds);
        a_vgs = a_vgd + a_vds;
      }
      bool l;
      if (a_vds >= 0) {
        a_vbs = pnjlim(a_vbs, acc.vbs, vt, vcritS, &l);
      } else {
        bool l2;
        double a_vbd = pnjlim(a_vbs - a_vds, acc.vbs - acc.vds, vt, vcritD, &l2);
        a_vbs = a_vbd + a_vds;
        l = l2;
      }
      lim = l;
    }
    double a_vbd = a_vbs - a_vds, a_vgd = a_vgs - a_vds;
    double a_vgb2 = a_vgs - a_vbs;
    vgs = a_vgs; vds = a_vds; vbs = a_vbs; vbd = a_vbd; vgd = a_vgd; vgb = a_vgb2;
    acc.vgs = a_vgs; acc.vds = a_vds; acc.vbs = a_vbs;
    bulkDiode(vbs, iss, &cbs, &gbs);
    bulkDiode(vbd, isd, &cbd, &gbd);
    channel(vgs, vds, vbs, &mode, &cdrain, &gm, &gds, &gmbs, &von, &vdsat);
    acc.von = von;
    tran = false;
    ag = 0;
    for (int k = 0; k < 5; k++) iq[k] = 0;
    dv[0] = vgs; dv[1] = vgd; dv[2] = vgb; dv[3] = vbs; dv[4] = vbd;
    if (lc.mode == ModeTran && (hasGcap || hasJcap)) {
      tran = true;
      ag = lc.method == Trapezoidal ? 2 / lc.H : 1 / lc.H;
      double mc[3];
      meyerCaps(vgs, vgd, von, vdsat, mode, mc);
      double ov[3] = {cgso, cgdo, cgbo};
      for (int k = 0; k < 3; k++) capG[k] = (mc[k] + meyerAcc[k]) / 2 + ov[k];
      double qbs, cbsc, qbd, cbdc;
      bulkCharge(vbs, czbs, czbssw, &qbs, &cbsc);
      bulkCharge(vbd, czbd, czbdsw, &qbd, &cbdc);
      capB[0] = cbsc; capB[1] = cbdc;
      double qs[5] = {qPrev[0] + capG[0] * (vgs - vPrev[0]), qPrev[1] + capG[1] * (vgd - vPrev[1]),
                      qPrev[2] + capG[2] * (vgb - vPrev[2]), qbs, qbd};
      for (int k = 0; k < 5; k++) {
        iq[k] = ag * (qs[k] - qPrev[k]);
        if (lc.method == Trapezoidal) iq[k] -= iqPrev[k];
      }
    }
    stamp(lc);
  }
  bool nonlinear() const override { return true; }
  bool limited() const override { return lim; }
  bool convTest(const double* x, int nx, double reltol, double abstol) override {  // :618-631
    double a_vgs, a_vds, a_vbs, a_vgb;
    voltages(x, nx, &a_vgs, &a_vds, &a_vbs, &a_vgb);
    double a_vbd = a_vbs - a_vds;
    double i0, gc[3];
    channelBranch(&i0, gc);
    double ich = i0 + gc[0] * (a_vgs - vgs) + gc[1] * (a_vds - vds) + gc[2] * (a_vbs - vbs);
    double ibs = cbs + gbs * (a_vbs - vbs);
    double ibd = cbd + gbd * (a_vbd - vbd);
    double md, cd, t1, t2, t3, t4, t5;
    channel(a_vgs, a_vds, a_vbs, &md, &cd, &t1, &t2, &t3, &t4, &t5);
    double ichN = md * cd;
    double ibsN, ibdN, gg;
    bulkDiode(a_vbs, iss, &ibsN, &gg);
    bulkDiode(a_vbd, isd, &ibdN, &gg);
    auto ok = [&](double a, double bb) { return std::fabs(a - bb) <= reltol * std::max(std::fabs(a), std::fabs(bb)) + abstol; };
    return ok(ich - ibd, ichN - ibdN) && ok(ibs + ibd, ibsN + ibdN);
  }
  void terminalCurrents(const double* x, int nx, double* id, double* ig, double* is, double* ib, double* ps) {  // :635-657
    double a_vgs, a_vds, a_vbs, a_vgb;
    voltages(x, nx, &a_vgs, &a_vds, &a_vbs, &a_vgb);
    double a_vbd = a_vbs - a_vds, a_vgd = a_vgs - a_vds;
    double t = typ;
    double i0, gc[3];
    channelBranch(&i0, gc);
    double ich = t * (i0 + gc[0] * (a_vgs - vgs) + gc[1] * (a_vds - vds) + gc[2] * (a_vbs - vbs));
    double ibs = t * (cbs + gbs * (a_vbs - vbs));
    double ibd = t * (cbd + gbd * (a_vbd - vbd));
    double iqv[5] = {0, 0, 0, 0, 0};
    *ps = 0;
    if (tran) {
      double v[5] = {a_vgs, a_vgd, a_vgb, a_vbs, a_vbd};
      for (int k = 0; k < 5; k++) {
        iqv[k] = t * (iq[k] + ag * capk(k) * (v[k] - dv[k]));
        *ps += v[k] * t * iqv[k];
      }
    }
    *id = ich - ibd - iqv[1] - iqv[4];
    *is = -ich - ibs - iqv[0] - iqv[3];
    *ig = iqv[0] + iqv[1] + iqv[2];
    *ib = ibs + ibd - iqv[2] + iqv[3] + iqv[4];
  }
  OperatingPoint operating(const double* x, int nx) override {  // :660-665
    double id, ig, is, ib, ps;
    terminalCurrents(x, nx, &id, &ig, &is, &ib, &ps);
    double vd = at(x, nx, u[0]), vs = at(x, nx, u[1]);
    double p = vd * id + at(x, nx, g) * ig + vs * is + at(x, nx, b) * ib;
    OperatingPoint op;
    op.voltage = vd - vs;
    op.current = id;
    op.power = p;
    op.storage = ps;
    return op;
  }
  void accept(const double* x, int nx, double t, int md_mode) override {  // :668-705
    double a_vgs, a_vds, a_vbs, a_vgb;
    voltages(x, nx, &a_vgs, &a_vds, &a_vbs, &a_vgb);
    double a_vgd = a_vgs - a_vds, a_vbd = a_vbs - a_vds;
    double md, cd, t1, t2, t3, v_on, v_dsat;
    channel(a_vgs, a_vds, a_vbs, &md, &cd, &t1, &t2, &t3, &v_on, &v_dsat);
    acc.vgs = a_vgs; acc.vds = a_vds; acc.vbs = a_vbs; acc.von = v_on;
    acc.vgd = a_vgd; acc.vgb = a_vgb; acc.vbd = a_vbd;
    double mc[3];
    meyerCaps(a_vgs, a_vgd, v_on, v_dsat, md, mc);
    double v[5] = {a_vgs, a_vgd, a_vgb, a_vbs, a_vbd};
    double qbs, qbd, cc;
    bulkCharge(a_vbs, czbs, czbssw, &qbs, &cc);
    bulkCharge(a_vbd, czbd, czbdsw, &qbd, &cc);
    if (md_mode == ModeTran && tran) {
      double qs[5];
      for (int k = 0; k < 3; k++) qs[k] = qPrev[k] + capG[k] * (v[k] - vPrev[k]);
      qs[3] = qbs;
      qs[4] = qbd;
      for (int k = 0; k < 5; k++) {
        iqPrev[k] = iq[k] + ag * capk(k) * (v[k] - dv[k]);
        qPrev[k] = qs[k];
        hist[k].push(t, qs[k]);
      }
    } else {
      double ov[3] = {cgso, cgdo, cgbo};
      for (int k = 0; k < 3; k++) qPrev[k] = (mc[k] + ov[k]) * v[k];
      qPrev[3] = qbs;
      qPrev[4] = qbd;
      for (int k = 0; k < 5; k++) {
        iqPrev[k] = 0;
        hist[k].reset();
        hist[k].push(t, qPrev[k]);
      }
    }
    for (int k = 0; k < 5; k++) vPrev[k] = v[k];
    for (int k = 0; k < 3; k++) meyerAcc[k] = mc[k];
  }
  bool reactive() const override { return true; }
  double truncate(const double* x, int nx, const TruncContext& tc) override {  // :734-757
    if (!tran) return kInf;
    double a_vgs, a_vds, a_vbs, a_vgb;
    voltages(x, nx, &a_vgs, &a_vds, &a_vbs, &a_vgb);
    double v[5] = {a_vgs, a_vgs - a_vds, a_vgb, a_vbs, a_vbs - a_vds};
    double h = kInf;
    for (int k = 0; k < 5; k++) {
      double qk, cc;
      if (k < 3) qk = qPrev[k] + capG[k] * (v[k] - vPrev[k]);
      else if (k == 3) bulkCharge(v[k], czbs, czbssw, &qk, &cc);
      else bulkCharge(v[k], czbd, czbdsw, &qk, &cc);
      if (qk == 0 && qPrev[k] == 0) continue;
      double i = iq[k] + ag * capk(k) * (v[k] - dv[k]);
      h = std::min(h, lteStep(hist[k], qk, i, iqPrev[k], tc.absTol, tc));
    }
    return h;
  }
  // StoredEnergy, device/mosfet.go:707-728: the junction charges' co-energy
  // by quadrature on 8 panels, ½·(C_meyer + C_overlap)·v² of the gate charges
  double storedEnergy(const double* x, int nx) override {
    if (!hasGcap && !hasJcap) return 0;
    double a_vgs, a_vds, a_vbs, a_vgb;
    voltages(x, nx, &a_vgs, &a_vds, &a_vbs, &a_vgb);
    double v[5] = {a_vgs, a_vgs - a_vds, a_vgb, a_vbs, a_vbs - a_vds};
    double w = 0.0;
    double cz[2][2] = {{czbs, czbssw}, {czbd, czbdsw}};
    for (int k = 0; k < 2; k++) {
      double vv = v[3 + k];
      if (vv == 0 || cz[k][0] + cz[k][1] == 0) continue;
      auto qf = [&](double y) {
        double qv, cc;
        bulkCharge(y, cz[k][0], cz[k][1], &qv, &cc);
        return qv;
      };
      w += vv * qf(vv) - glIntegrate(qf, 0.0, vv, 8);
    }
    double md, cd, t1, t2, t3, v_on, v_dsat;
    channel(v[0], v[0] - v[1], v[3], &md, &cd, &t1, &t2, &t3, &v_on, &v_dsat);
    double mc[3];
    meyerCaps(v[0], v[1], v_on, v_dsat, md, mc);
    double ov[3] = {cgso, cgdo, cgbo};
    for (int k = 0; k < 3; k++) w += 0.5 * (mc[k] + ov[k]) * v[k] * v[k];
    return w;
  }
  void ports(const OperatingPoint& op, Accumulator& accu) override { accu.addPort(2, -(op.power - op.storage)); }
  void loadAC(ACContext& ctx) override;
};

// Gummel–Poon BJT, device/bjt.go:14-765 (transport and base currents,
// base-resistance modulation, transit-time and depletion charges, substrate
// junction, pnjlim limiting, LoadAC).
struct Bjt : Device {
  // Parameters (area-scaled) and derived constants, NewBJT :142-242.
  double typ = 1, IS = 1e-16, BF = 100, NF = 1, VAF = 0, IKF = 0, ISE = 0, NE = 1.5, BR = 1, NR = 1, VAR = 0, IKR = 0;
  double ISC = 0, NC = 2, RB = 0, IRB = 0, RBM = 0, RE = 0, RC = 0, CJE = 0, VJE = 0.75, MJE = 0.33, TF = 0, XTF = 0;
  double VTF = 0, ITF = 0, CJC = 0, VJC = 0.75, MJC = 0.33, XCJC = 1, TR = 0, CJS = 0, VJS = 0.75, MJS = 0, FC = 0.5;
  double ISS = 0, NS = 1, SUBS = 1;
  double vt = 0, vcrit = 0, rbpr = 0, rbpi = 0, gcpr = 0, gepr = 0, ovtf = 0, f1e = 0, f2e = 0, f3e = 0, f1c = 0,
         f2c = 0, f3c = 0;
  bool hasCharge = false;
  int b = -1, s = -1, cp = -1, bp = -1, ep = -1, sa = -1, sb = -1;

  struct DC {  // bjtDC, :118-127
    double cbe = 0, gbe = 0, cbc = 0, gbc = 0, it = 0, itBE = 0, itBC = 0, ibe = 0, gpi = 0, ibc = 0, gmu = 0;
    double qb = 0, dqbdve = 0, dqbdvc = 0, gx = 0;
  };
  struct Charge {  // bjtCharge, :130-134
    double q[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0}, dqbeDvbc = 0;
  };
  // Linearization of the most recent Load.
  double vbe = 0, vbc = 0, vbx = 0, vcs = 0, isub = 0, gsub = 0, ag = 0, iq[4] = {0, 0, 0, 0}, gmin = 0;
  DC dc;
  Charge ch;
  bool lim = false, tran = false;
  // Accepted state.
  double vbeAcc = 0, vbcAcc = 0, vbxAcc = 0, vcsAcc = 0, qPrev[4] = {0, 0, 0, 0}, iqPrev[4] = {0, 0, 0, 0};
  History hist[4];
  double omega = 0;
  struct { double gx = 0, itBE = 0, itBC = 0, gpi = 0, gmu = 0; cplx ybe, ybeBC, ybc, ybx, ycs; } ac;

  void setup(SetupContext& c) override {  // :250-268
    setupNodes(c);
    b = c.unknown(nodes[2]);
    s = c.unknown(nodes[3]);
    cp = u[0]; bp = b; ep = u[1];
    if (RC > 0) cp = c.allocUnknown();
    if (RB > 0) bp = c.allocUnknown();
    if (RE > 0) ep = c.allocUnknown();
    sa = s; sb = cp;
    if (SUBS < 0) { sa = bp; sb = s; }
  }
  // junction, :272-283: IS(e^(v/nvt) - 1) with the cubic reverse tail below -3 nvt.
  static void junction(double is, double v, double nvt, double* i, double* g) {
    if (is == 0) { *i = 0; *g = 0; return; }
    if (v >= -3 * nvt) {
      double e = std::exp(std::min(v / nvt, 700.0));
      *i = is * (e - 1);
      *g = is * e / nvt;
      return;
    }
    double a = 3 * nvt / (v * kE);
    a = a * a * a;
    *i = -is * (1 + a);
    *g = is * 3 * a / v;
  }
  void substrate(double v, double* i, double* g) const {  // :288-291
    junction(ISS, v, NS * vt, i, g);
    *i += gmin * v;
    *g += gmin;
  }
  DC evalDC(double a_vbe, double a_vbc) const {  // :294-355
    DC d;
    junction(IS, a_vbe, NF * vt, &d.cbe, &d.gbe);
    double cben, gben, cbcn, gbcn;
    junction(ISE, a_vbe, NE * vt, &cben, &gben);
    cben += gmin * a_vbe;
    gben += gmin;
    junction(IS, a_vbc, NR * vt, &d.cbc, &d.gbc);
    junction(ISC, a_vbc, NC * vt, &cbcn, &gbcn);
    cbcn += gmin * a_vbc;
    gbcn += gmin;
    double ivaf = VAF > 0 ? 1 / VAF : 0, ivar = VAR > 0 ? 1 / VAR : 0;
    double oik = IKF > 0 ? 1 / IKF : 0, oikr = IKR > 0 ? 1 / IKR : 0;
    double q1 = 1 / (1 - ivaf * a_vbc - ivar * a_vbe);
    if (oik == 0 && oikr == 0) {
      d.qb = q1;
      d.dqbdve = q1 * d.qb * ivar;
      d.dqbdvc = q1 * d.qb * ivaf;
    } else {
      double q2 = oik * d.cbe + oikr * d.cbc;
      double arg = std::max(0.0, 1 + 4 * q2);
      double sq = arg != 0 ? std::sqrt(arg) : 1.0;
      d.qb = q1 * (1 + sq) / 2;
      d.dqbdve = q1 * (d.qb * ivar + oik * d.gbe / sq);
      d.dqbdvc = q1 * (d.qb * ivaf + oikr * d.gbc / sq);
    }
    d.it = (d.cbe - d.cbc) / d.qb;
    d.itBE = (d.gbe - d.it * d.dqbdve) / d.qb;
    d.itBC = (-d.gbc - d.it * d.dqbdvc) / d.qb;
    d.ibe = d.cbe / BF + cben;
    d.gpi = d.gbe / BF + gben;
    d.ibc = d.cbc / BR + cbcn;
    d.gmu = d.gbc / BR + gbcn;
    double gx = rbpr + rbpi / d.qb;
    if (IRB > 0 && rbpi != 0) {  // current crowding
      double cb = d.ibe + d.ibc;
      double a1 = std::max(cb / IRB, 1e-9);
      double z = (-1 + std::sqrt(1 + 14.59025 * a1)) / 2.4317 / std::sqrt(a1);
      double tz = std::tan(z);
      gx = rbpr + 3 * rbpi * (tz - z) / z / tz / tz;
    }
    if (gx != 0) gx = 1 / gx;
    d.gx = gx;
    return d;
  }
  // depletionExp: the BJT's depletion charge (:358-370), arg^-m as exp(-m log arg).
  Charge charges(double a_vbe, double a_vbc, double a_vbx, double a_vcs, const DC& d) const {  // :373-410
    Charge c;
    double cbe = d.cbe, gbe = d.gbe;
    if (TF != 0 && a_vbe > 0) {
      double argtf = 0, arg2 = 0, arg3 = 0;
      if (XTF != 0) {
        argtf = XTF;
        if (ovtf != 0) argtf *= std::exp(a_vbc * ovtf);
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
      c.dqbeDvbc = TF * (arg3 - cbe * d.dqbdvc) / d.qb;
    }
    double qd, cd;
    depletion(a_vbe, CJE, VJE, MJE, FC, f1e, f2e, f3e, &qd, &cd);
    c.q[0] = TF * cbe + qd;
    c.c[0] = TF * gbe + cd;
    double czbc = CJC * XCJC;
    depletion(a_vbc, czbc, VJC, MJC, FC, f1c, f2c, f3c, &qd, &cd);
    c.q[1] = TR * d.cbc + qd;
    c.c[1] = TR * d.gbc + cd;
    depletion(a_vbx, CJC - czbc, VJC, MJC, FC, f1c, f2c, f3c, &c.q[2], &c.c[2]);
    if (CJS != 0) {
      if (a_vcs < 0) {
        double arg = 1 - a_vcs / VJS;
        double sarg = std::exp(-MJS * std::log(arg));
        c.q[3] = VJS * CJS * (1 - arg * sarg) / (1 - MJS);
        c.c[3] = CJS * sarg;
      } else {
        c.q[3] = a_vcs * CJS * (1 + MJS * a_vcs / (2 * VJS));
        c.c[3] = CJS * (1 + MJS * a_vcs / VJS);
      }
    }
    return c;
  }
  void voltages(const double* x, int nx, double v[4]) const {  // :413-418
    double vb = at(x, nx, b), vbp = at(x, nx, bp), vcp = at(x, nx, cp), vep = at(x, nx, ep);
    v[0] = typ * (vbp - vep);
    v[1] = typ * (vbp - vcp);
    v[2] = typ * (vb - vcp);
    v[3] = typ * (at(x, nx, sa) - at(x, nx, sb));
  }
  // branch, :447-469: current i0 + g1 dv1 + g2 dv2 from a to b.
  void branch(LoadContext& lc, int a, int bb, double i0, double g1, int p1, int n1, double v10, double g2, int p2, int n2,
              double v20) {
    Stamper* m = lc.M;
    if (g1 != 0) {
      m->add(a, p1, g1); m->add(a, n1, -g1); m->add(bb, p1, -g1); m->add(bb, n1, g1);
    }
    if (g2 != 0) {
      m->add(a, p2, g2); m->add(a, n2, -g2); m->add(bb, p2, -g2); m->add(bb, n2, g2);
    }
    double ieq = typ * (i0 - g1 * v10 - g2 * v20);
    lc.addRHS(a, -ieq);
    lc.addRHS(bb, ieq);
  }
  void stamp(LoadContext& lc) {  // :492-515
    const DC& d = dc;
    branch(lc, cp, ep, d.it, d.itBE, bp, ep, vbe, d.itBC, bp, cp, vbc);
    branch(lc, bp, ep, d.ibe, d.gpi, bp, ep, vbe, 0, 0, 0, 0);
    branch(lc, bp, cp, d.ibc, d.gmu, bp, cp, vbc, 0, 0, 0, 0);
    branch(lc, sa, sb, isub, gsub, sa, sb, vcs, 0, 0, 0, 0);
    if (tran) {
      branch(lc, bp, ep, iq[0], ag * ch.c[0], bp, ep, vbe, ag * ch.dqbeDvbc, bp, cp, vbc);
      branch(lc, bp, cp, iq[1], ag * ch.c[1], bp, cp, vbc, 0, 0, 0, 0);
      branch(lc, b, cp, iq[2], ag * ch.c[2], b, cp, vbx, 0, 0, 0, 0);
      branch(lc, sa, sb, iq[3], ag * ch.c[3], sa, sb, vcs, 0, 0, 0, 0);
    }
    if (bp != b) Mosfet::stampG(lc.M, b, bp, d.gx);
    if (cp != u[0]) Mosfet::stampG(lc.M, u[0], cp, gcpr);
    if (ep != u[1]) Mosfet::stampG(lc.M, u[1], ep, gepr);
  }
  void load(LoadContext& lc) override {  // :421-444
    gmin = lc.gmin;
    double v[4];
    voltages(lc.X, lc.nx, v);
    lim = false;
    if (lc.initJct) {
      v[0] = vcrit; v[1] = 0; v[2] = 0; v[3] = 0;
      lim = true;
    } else {
      bool l1, l2;
      v[0] = pnjlim(v[0], vbeAcc, vt, vcrit, &l1);
      v[1] = pnjlim(v[1], vbcAcc, vt, vcrit, &l2);
      lim = l1 || l2;
    }
    vbeAcc = v[0]; vbcAcc = v[1];
    vbe = v[0]; vbc = v[1]; vbx = v[2]; vcs = v[3];
    dc = evalDC(vbe, vbc);
    substrate(vcs, &isub, &gsub);
    tran = false;
    ag = 0;
    for (double& q : iq) q = 0;
    if (lc.mode == ModeTran && hasCharge) {
      ch = charges(vbe, vbc, vbx, vcs, dc);
      tran = true;
      if (lc.method == Trapezoidal) {
        ag = 2 / lc.H;
        for (int k = 0; k < 4; k++) iq[k] = ag * (ch.q[k] - qPrev[k]) - iqPrev[k];
      } else {
        ag = 1 / lc.H;
        for (int k = 0; k < 4; k++) iq[k] = ag * (ch.q[k] - qPrev[k]);
      }
    }
    stamp(lc);
  }
  bool nonlinear() const override { return true; }
  bool limited() const override { return lim; }
  bool convTest(const double* x, int nx, double reltol, double abstol) override {  // :521-532
    double v[4];
    voltages(x, nx, v);
    const DC& d = dc;
    double dbe = v[0] - vbe, dbc = v[1] - vbc;
    double ichat = d.it - d.ibc + d.itBE * dbe + (d.itBC - d.gmu) * dbc;
    double ibhat = d.ibe + d.ibc + d.gpi * dbe + d.gmu * dbc;
    DC nw = evalDC(v[0], v[1]);
    double ic = nw.it - nw.ibc, ib = nw.ibe + nw.ibc;
    return std::fabs(ichat - ic) <= reltol * std::max(std::fabs(ichat), std::fabs(ic)) + abstol &&
           std::fabs(ibhat - ib) <= reltol * std::max(std::fabs(ibhat), std::fabs(ib)) + abstol;
  }
  // terminalCurrents, :537-570: currents into c, b, e, s from the linearization.
  void terminalCurrents(const double* x, int nx, double* ic, double* ib, double* ie, double* is, double* ps) const {
    double v[4];
    voltages(x, nx, v);
    double t = typ;
    const DC& d = dc;
    double dbe = v[0] - vbe, dbc = v[1] - vbc;
    double it = t * (d.it + d.itBE * dbe + d.itBC * dbc);
    double ibe = t * (d.ibe + d.gpi * dbe);
    double ibc = t * (d.ibc + d.gmu * dbc);
    double iqbe = 0, iqbc = 0, iqbx = 0, iqcs = 0;
    *ps = 0;
    if (tran) {
      iqbe = t * (iq[0] + ag * ch.c[0] * dbe + ag * ch.dqbeDvbc * dbc);
      iqbc = t * (iq[1] + ag * ch.c[1] * dbc);
      iqbx = t * (iq[2] + ag * ch.c[2] * (v[2] - vbx));
      iqcs = t * (iq[3] + ag * ch.c[3] * (v[3] - vcs));
      *ps = v[0] * t * iqbe + v[1] * t * iqbc + v[2] * t * iqbx + v[3] * t * iqcs;
    }
    double isb = t * (isub + gsub * (v[3] - vcs)) + iqcs;
    *ic = it - ibc - iqbc - iqbx;
    *ib = ibe + ibc + iqbe + iqbc + iqbx;
    *ie = -it - ibe - iqbe;
    if (SUBS > 0) {
      *is = isb;
      *ic = *ic - isb;
    } else {
      *ib = *ib + isb;
      *is = -isb;
    }
  }
  OperatingPoint operating(const double* x, int nx) override {  // :574-579
    double ic, ib, ie, is, ps;
    terminalCurrents(x, nx, &ic, &ib, &ie, &is, &ps);
    double vc = at(x, nx, u[0]), ve = at(x, nx, u[1]);
    OperatingPoint op;
    op.power = vc * ic + at(x, nx, b) * ib + ve * ie + at(x, nx, s) * is;
    op.voltage = vc - ve;
    op.current = ic;
    op.storage = ps;
    return op;
  }
  void accept(const double* x, int nx, double t, int mode) override {  // :582-605
    double v[4];
    voltages(x, nx, v);
    vbeAcc = v[0]; vbcAcc = v[1]; vbxAcc = v[2]; vcsAcc = v[3];
    DC d = evalDC(v[0], v[1]);
    Charge c = charges(v[0], v[1], v[2], v[3], d);
    double dv[4] = {v[0] - vbe, v[1] - vbc, v[2] - vbx, v[3] - vcs};
    for (int k = 0; k < 4; k++) {
      if (mode == ModeTran && tran) {
        double i = iq[k] + ag * ch.c[k] * dv[k];
        if (k == 0) i += ag * ch.dqbeDvbc * (v[1] - vbc);
        iqPrev[k] = i;
      } else {
        iqPrev[k] = 0;
      }
      if (mode != ModeTran) hist[k].reset();
      qPrev[k] = c.q[k];
      hist[k].push(t, c.q[k]);
    }
  }
  bool reactive() const override { return true; }
  double truncate(const double* x, int nx, const TruncContext& tc) override {  // :655-675
    if (!hasCharge || !tran) return kInf;
    double v[4];
    voltages(x, nx, v);
    DC d = evalDC(v[0], v[1]);
    Charge c = charges(v[0], v[1], v[2], v[3], d);
    double dv[4] = {v[0] - vbe, v[1] - vbc, v[2] - vbx, v[3] - vcs};
    double h = kInf;
    for (int k = 0; k < 4; k++) {
      if (c.q[k] == 0 && qPrev[k] == 0) continue;
      double i = iq[k] + ag * ch.c[k] * dv[k];
      if (k == 0) i += ag * ch.dqbeDvbc * dv[1];
      h = std::min(h, lteStep(hist[k], c.q[k], i, iqPrev[k], tc.absTol, tc));
    }
    return h;
  }
  // StoredEnergy, device/bjt.go:624-648: Σ_k v_k·q_k(v_k) − ∫₀^{v_k} q_k, each
  // charge along its own voltage with the others held, quadrature on 8 panels
  double storedEnergy(const double* x, int nx) override {
    if (!hasCharge) return 0;
    double v[4];
    voltages(x, nx, v);
    double w = 0.0;
    for (int k = 0; k < 4; k++) {
      if (v[k] == 0) continue;
      auto qk = [&](double y) {
        double vv[4] = {v[0], v[1], v[2], v[3]};
        vv[k] = y;
        DC d = evalDC(vv[0], vv[1]);
        return charges(vv[0], vv[1], vv[2], vv[3], d).q[k];
      };
      w += v[k] * qk(v[k]) - glIntegrate(qk, 0.0, v[k], 8);
    }
    return w;
  }
  void ports(const OperatingPoint& op, Accumulator& accu) override { accu.addPort(2, -(op.power - op.storage)); }
  // acBranch, :718-731
  static void acBranch(CStamper* m, int a, int bb, cplx y1, int p1, int n1, cplx y2, int p2, int n2) {
    if (y1 != cplx(0)) {
      m->add(a, p1, y1); m->add(a, n1, -y1); m->add(bb, p1, -y1); m->add(bb, n1, y1);
    }
    if (y2 != cplx(0)) {
      m->add(a, p2, y2); m->add(a, n2, -y2); m->add(bb, p2, -y2); m->add(bb, n2, y2);
    }
  }
  void loadAC(ACContext& ctx) override {  // :690-716
    DC d = evalDC(vbeAcc, vbcAcc);
    Charge c = charges(vbeAcc, vbcAcc, vbxAcc, vcsAcc, d);
    double w = ctx.omega;
    omega = w;
    ac.itBE = d.itBE; ac.itBC = d.itBC; ac.gpi = d.gpi; ac.gmu = d.gmu; ac.gx = d.gx;
    ac.ybe = cplx(0, w * c.c[0]);
    ac.ybeBC = cplx(0, w * c.dqbeDvbc);
    double is_, gs;
    substrate(vcsAcc, &is_, &gs);
    ac.ybc = cplx(0, w * c.c[1]);
    ac.ybx = cplx(0, w * c.c[2]);
    ac.ycs = cplx(gs, w * c.c[3]);
    CStamper* m = ctx.M;
    acBranch(m, cp, ep, cplx(ac.itBE, 0), bp, ep, cplx(ac.itBC, 0), bp, cp);
    acBranch(m, bp, ep, cplx(ac.gpi, 0) + ac.ybe, bp, ep, ac.ybeBC, bp, cp);
    acBranch(m, bp, cp, cplx(ac.gmu, 0) + ac.ybc, bp, cp, 0, 0, 0);
    acBranch(m, b, cp, ac.ybx, b, cp, 0, 0, 0);
    acBranch(m, sa, sb, ac.ycs, sa, sb, 0, 0, 0);
    if (bp != b) stampY(m, b, bp, cplx(ac.gx, 0));
    if (cp != u[0]) stampY(m, u[0], cp, cplx(gcpr, 0));
    if (ep != u[1]) stampY(m, u[1], ep, cplx(gepr, 0));
  }
  // ACPower, :759-765 (acCurrents :734-751): half the sum of V_k conj(I_k).
  cplx acPower(const cplx* x, int nx, cplx* ic_out = nullptr) const {
    auto cat = [&](int i) { return (i < 0 || i >= nx) ? cplx(0) : x[i]; };
    cplx vbe_ = cat(bp) - cat(ep), vbc_ = cat(bp) - cat(cp), vbx_ = cat(b) - cat(cp), vcs_ = cat(sa) - cat(sb);
    cplx it = cplx(ac.itBE, 0) * vbe_ + cplx(ac.itBC, 0) * vbc_;
    cplx ibe = (cplx(ac.gpi, 0) + ac.ybe) * vbe_ + ac.ybeBC * vbc_;
    cplx ibc = (cplx(ac.gmu, 0) + ac.ybc) * vbc_;
    cplx ix = ac.ybx * vbx_, isb = ac.ycs * vcs_;
    cplx ic, ib, ie = -it - ibe, is;
    if (SUBS > 0) {
      ic = it - ibc - ix - isb; ib = ibe + ibc + ix; is = isb;
    } else {
      ic = it - ibc - ix; ib = ibe + ibc + ix + isb; is = -isb;
    }
    cplx sum = cat(u[0]) * std::conj(ic) + cat(b) * std::conj(ib) + cat(u[1]) * std::conj(ie) + cat(s) * std::conj(is);
    if (ic_out) *ic_out = ic;  // OperatingAC's current, :767-771
    return cplx(sum.real() / 2, sum.imag() / 2);
  }
};

}  // namespace oracle
