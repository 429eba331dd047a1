// kh_report.h — one element's entry of output.Result (output/result.go:15-35)
// at the present solution and device state, as analysis's Sim.result builds
// it (analysis/op.go:361-404): device.Operating's voltage, current, absorbed
// and storage power; device.Reactive.StoredEnergy; device.ModelSupplier's
// model-supplied power; and the power through the element's modeled exterior
// port (device.ExteriorPower), whose channel is fixed by its kind.
#pragma once
#include "kh_devices.h"

namespace kh {

// Fields of one element's report, in this order.
enum {
  RP_VOLTAGE = 0,   // V(n+) - V(n-)
  RP_CURRENT,       // n+ -> n-
  RP_POWER,         // absorbed
  RP_STORAGE,       // part of the absorbed power going into storage
  RP_ENERGY,        // stored energy [J]
  RP_SUPPLIED,      // model-supplied power (negative resistances, controlled sources)
  RP_EXTERIOR,      // power through the modeled exterior port: + in from the exterior
  RP_CHANNEL,       // that port's channel: 0 electrical, 2 thermal, -1 none
  RP_FIELDS
};

// 16-point Gauss–Legendre nodes and weights on [-1, 1] (Abramowitz & Stegun,
// Table 25.4), to the 16 significant digits the reference's table carries
// (device/bjt.go:603-608), so the quadratures below round alike.
constexpr double kGL16x[8] = {0.0950125098376374, 0.2816035507792589, 0.4580167776572274, 0.6178762444026438,
                              0.7554044083550030, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499};
constexpr double kGL16w[8] = {0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                              0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541};

// ∫_a^b f by 16-point Gauss–Legendre on n equal panels, summed panel by
// panel and node pair by node pair (device/bjt.go:611-622).
template <class F>
KH_HD inline double gl_integrate(F f, double a, double b, int n) {
  const double h = (b - a) / (double)n;
  double s = 0.0;
  for (int p = 0; p < n; p++) {
    const double mid = a + ((double)p + 0.5) * h, half = h / 2;
    for (int k = 0; k < 8; k++) s += kGL16w[k] * (f(mid - half * kGL16x[k]) + f(mid + half * kGL16x[k]));
  }
  return s * h / 2;
}

// Diode: ∫_0^vd v·C(v) dv = vd·q(vd) − ∫_0^vd q(u) du in closed form, region by
// region (device/diode.go:329-383), at the junction voltage of the present
// solution (the internal anode when RS > 0).
KH_HD inline double diode_int_depletion(const Lane& L, const int32_t* e, double v) {
  const double CJO = L.P(e, 3), VJ = L.P(e, 4), M = L.P(e, 5), FC = L.P(e, 9);
  if (CJO == 0) return 0;
  const double fcv = FC * VJ;
  auto below = [&](double u) { return CJO * VJ / (1 - M) * (u + VJ / (2 - M) * (pow(1 - u / VJ, 2 - M) - 1)); };
  if (v < fcv) return below(v);
  const double f1 = VJ * (1 - pow(1 - FC, 1 - M)) / (1 - M);
  const double f2 = pow(1 - FC, 1 + M);
  const double f3 = 1 - FC * (1 + M);
  const double dv = v - fcv;
  return below(fcv) + CJO * f1 * dv +
         CJO / f2 * (f3 * dv * dv / 2 + M / (2 * VJ) * ((v * v * v - fcv * fcv * fcv) / 3 - fcv * fcv * dv));
}
KH_HD inline double diode_int_current(const Lane& L, const int32_t* e, double v) {
  const double IS = L.P(e, 0), vte = L.P(e, 11), xbv = L.P(e, 13);
  const bool hasBV = L.P(e, 10) != 0;
  auto fwd = [&](double u) { return IS * (vte * expm1(fmin(u / vte, 700.0)) - u); };
  const double v1 = -3 * vte;
  if (v >= v1) return fwd(v);
  double total = fwd(v1);
  const double a3 = pow(3 * vte / kE, 3.0);
  auto rev = [&](double u) { return -IS * (u - a3 / (2 * u * u)); };
  double v2 = v;
  if (hasBV && v < -xbv) v2 = -xbv;
  total += rev(v2) - rev(v1);
  if (v2 == v) return total;
  auto brk = [&](double u) { return IS * vte * exp(fmin(-(xbv + u) / vte, 700.0)); };
  return total + brk(v) - brk(v2);
}
KH_HD inline double stored_D(const Lane& L, const int32_t* e, const double* x) {
  if (L.P(e, 15) == 0) return 0;
  const double vd = L.xat(x, L.T(e, 2)) - L.xat(x, L.T(e, 1));
  double id, gd, q, cc;
  diode_current(L, e, vd, &id, &gd);
  diode_charge(L, e, vd, id, gd, &q, &cc);
  return vd * q - diode_int_depletion(L, e, vd) - L.P(e, 6) * diode_int_current(L, e, vd);
}

// MOSFET: the junction charges' co-energy by quadrature on 8 panels, plus
// ½·(C_meyer + C_overlap)·v² of the gate capacitances (device/mosfet.go:707-728).
KH_HD inline double stored_M(const Lane& L, const int32_t* e, const double* x) {
  if (L.P(e, 33) == 0 && L.P(e, 32) == 0) return 0;
  double vgs, vds, vbs, vgb;
  mos_voltages(L, e, x, &vgs, &vds, &vbs, &vgb);
  const double v[5] = {vgs, vgs - vds, vgb, vbs, vbs - vds};
  double w = 0.0;
  const double cz[2][2] = {{L.P(e, 15), L.P(e, 16)}, {L.P(e, 13), L.P(e, 14)}};
  for (int k = 0; k < 2; k++) {
    const double vv = v[3 + k];
    if (vv == 0 || cz[k][0] + cz[k][1] == 0) continue;
    auto qf = [&](double u) {
      double qv, c;
      mos_bulk_charge(L, e, u, cz[k][0], cz[k][1], &qv, &c);
      return qv;
    };
    w += vv * qf(vv) - gl_integrate(qf, 0.0, vv, 8);
  }
  const MosIds ch = mos_channel(L, e, v[0], v[0] - v[1], v[3]);
  double mc[3];
  mos_meyer_caps(L, e, v[0], v[1], ch.von, ch.vdsat, ch.md, mc);
  const double ov[3] = {L.P(e, 27), L.P(e, 28), L.P(e, 29)};
  for (int k = 0; k < 3; k++) w += 0.5 * (mc[k] + ov[k]) * v[k] * v[k];
  return w;
}

// BJT: Σ_k v_k·q_k(v_k) − ∫_0^{v_k} q_k(u) du over the four charges, each
// along its own voltage with the others held, by quadrature on 8 panels
// (device/bjt.go:630-648).
KH_HD inline double stored_Q(const Lane& L, const int32_t* e, const double* x) {
  if (L.P(e, QP_HASQ) == 0) return 0;
  double v[4];
  bjt_voltages(L, e, x, v);
  const double gmin = L.S(e, QS_GMIN);
  double w = 0.0;
  for (int k = 0; k < 4; k++) {
    if (v[k] == 0) continue;
    auto qk = [&](double u) {
      double vv[4] = {v[0], v[1], v[2], v[3]};
      vv[k] = u;
      const BjtDC d = bjt_eval_dc(L, e, vv[0], vv[1], gmin);
      return bjt_charges(L, e, vv, d).q[k];
    };
    w += v[k] * qk(v[k]) - gl_integrate(qk, 0.0, v[k], 8);
  }
  return w;
}

// The element's report (RP_* fields) at solution x.
KH_HD inline void report_element(const Lane& L, const int32_t* e, const double* x, double* out) {
  const OpPt o = operating_point(L, e, x);
  double energy = 0, supplied = 0, exterior = 0, channel = -1;
  switch (e[ET_KIND]) {
    case K_R:  // device/passive.go:99-114: heat out of the thermal port for R > 0
      if (L.P(e, 0) > 0) {
        exterior = -o.power;
        channel = 2;
      } else {
        supplied = -o.power;
      }
      break;
    case K_C: energy = 0.5 * L.P(e, 0) * o.voltage * o.voltage; break;  // device/reactive.go:171-175
    case K_L: energy = 0.5 * L.P(e, 0) * o.current * o.current; break;  // device/reactive.go:339-342
    case K_K: energy = L.P(e, 0) * x[L.T(e, 0)] * x[L.T(e, 1)]; break;  // M·i1·i2, device/mutual.go:132-133
    case K_V:
    case K_I:  // device/source.go:123-129, :171-177: the source's power enters from the exterior
      exterior = -o.power;
      channel = 0;
      break;
    case K_D:
      if (KH_HAS(K_D)) energy = stored_D(L, e, x);
      exterior = -(o.power - o.storage);
      channel = 2;
      break;
    case K_M:
      if (KH_HAS(K_M)) energy = stored_M(L, e, x);
      exterior = -(o.power - o.storage);
      channel = 2;
      break;
    case K_Q:
      if (KH_HAS(K_Q)) energy = stored_Q(L, e, x);
      exterior = -(o.power - o.storage);
      channel = 2;
      break;
    case K_E:
    case K_F:
    case K_G:
    case K_H:  // device/controlled.go:132-136: all of its power is model-supplied
      supplied = -o.power;
      break;
  }
  out[RP_VOLTAGE] = o.voltage;
  out[RP_CURRENT] = o.current;
  out[RP_POWER] = o.power;
  out[RP_STORAGE] = o.storage;
  out[RP_ENERGY] = energy;
  out[RP_SUPPLIED] = supplied;
  out[RP_EXTERIOR] = exterior;
  out[RP_CHANNEL] = channel;
}

// Totals of one circuit's ledger, in this order (output.Ledger,
// output/result.go:58-90).
enum {
  LG_INFLOW = 0,
  LG_OUTFLOW,
  LG_DISSIPATION,
  LG_STORAGE_RATE,
  LG_MODEL_SUPPLIED,
  LG_HYPOTHETICAL,  // always 0: no hypothetical ports on these devices
  LG_RESIDUAL,
  LG_SCALE,
  LG_ROUNDOFF,
  LG_RELATIVE,      // Ledger.RelativeResidual (output/result.go:100-109)
  LG_FIELDS
};

// Elements that store energy (device.Reactive): their storage power enters
// the ledger.
KH_HD inline bool kind_stores(int kind) {
  return kind == K_C || kind == K_L || kind == K_D || kind == K_M || kind == K_Q || kind == K_K;
}

// em.BuildLedger (em/energy.go:89-119) over the element reports rep[nel][RP_FIELDS]
// in element order: each element's ports in channel order (only the one
// modeled exterior port can carry power), then its storage power, then its
// model-supplied power; the residual and the relative residual with the
// round-off floor.
KH_HD inline void fold_ledger(const int32_t* et, int nel, const double* rep, double roundoff, double* out) {
  double inflow = 0, outflow = 0, diss = 0, stor = 0, sup = 0, scale = 0;
  for (int k = 0; k < nel; k++) {
    const double* r = rep + (int64_t)k * RP_FIELDS;
    const double pw = r[RP_EXTERIOR];
    const int ch = (int)r[RP_CHANNEL];
    if (ch >= 0 && pw != 0) {  // Accumulator.AddPort
      scale += fabs(pw);
      if (ch == 2 && pw < 0) diss += -pw;
      else if (pw > 0) inflow += pw;
      else outflow += -pw;
    }
    if (kind_stores(et[(int64_t)k * ET_W + ET_KIND])) {  // AddStorage
      stor += r[RP_STORAGE];
      scale += fabs(r[RP_STORAGE]);
    }
    if (r[RP_SUPPLIED] != 0) {  // AddModelSupplied
      sup += r[RP_SUPPLIED];
      scale += fabs(r[RP_SUPPLIED]);
    }
  }
  const double res = inflow + sup - outflow - diss - stor;
  out[LG_INFLOW] = inflow;
  out[LG_OUTFLOW] = outflow;
  out[LG_DISSIPATION] = diss;
  out[LG_STORAGE_RATE] = stor;
  out[LG_MODEL_SUPPLIED] = sup;
  out[LG_HYPOTHETICAL] = 0;
  out[LG_RESIDUAL] = res;
  out[LG_SCALE] = scale;
  out[LG_ROUNDOFF] = roundoff;
  out[LG_RELATIVE] = (scale == 0 || fabs(res) <= fmax(1e-21, roundoff)) ? 0.0 : fabs(res / scale);
}

}  // namespace kh
