// kh_ac.h — small-signal AC systems of batched circuits (analysis/ac.go:76-193).
//
// After a circuit's operating point (and polish), every device is linearized
// at its accepted state. One AC system = one (circuit, frequency) pair: the
// devices' LoadAC stamps at omega = 2*pi*f go into complex slot values (the
// circuit's slot table, kh_symbolic.cpp, covers every pair LoadAC adds), the
// system is factored and solved with the complex Gilbert–Peierls LU
// (kh_lu.h with T = kcplx, as CSparse), and the complex-Tellegen ledger
// (the sum of the elements' absorbed complex power against its scale and the
// round-off floor) is evaluated from the solution.
//
// Systems are stored structure-of-arrays over the AC lanes l = f*B + c
// (consecutive lanes: consecutive circuits at one frequency).
#pragma once
#include "kh_devices.h"

namespace kh {

struct ACLane {
  Lane dev;            // circuit c's parameters and accepted state (per-circuit view, B = 1)
  int64_t l, BL;       // this system's lane and the number of AC lanes
  kcplx* val;          // [nslots][BL] stamped values (zeroed by the caller)
  kcplx* rhs;          // [n][BL]
  uint8_t* present;    // circuit c's [nslots] marks of stamped slots, or null
  double omega;

  KH_HD void add(const int32_t* e, int ti, int tj, kcplx v) {
    int s = e[ET_SLOT + ti * KH_NT + tj];
    if (s < 0) return;  // ground (solver/sparse.go:115-117)
    int64_t ix = (int64_t)s * BL + l;
    val[ix] = val[ix] + v;
    if (present) present[s] = 1;
  }
  // stampY, device/device.go:333-338
  KH_HD void stampY(const int32_t* e, int ta, int tb, kcplx y) {
    add(e, ta, ta, y);
    add(e, tb, tb, y);
    add(e, ta, tb, -y);
    add(e, tb, ta, -y);
  }
  KH_HD void add_rhs(int i, kcplx v) {
    if (i < 0 || i >= dev.n) return;
    int64_t ix = (int64_t)i * BL + l;
    rhs[ix] = rhs[ix] + v;
  }
  KH_HD kcplx xat(const kcplx* x, int i) const {
    return (i < 0 || i >= dev.n) ? kcplx{0, 0} : x[(int64_t)i * BL + l];
  }
};

// Source AC phasor: cmplx.Rect(mag, phase in degrees) (device/source.go:31-35).
KH_HD inline kcplx source_ac(const Lane& L, const int32_t* e) {
  double mag = L.P(e, 10), ph = L.P(e, 11) * 3.14159265358979323846 / 180;
  return {mag * cos(ph), mag * sin(ph)};
}

// MOSFET small-signal quantities at the accepted point (device/mosfet.go:779-814).
struct MosAC {
  double gch[3], gbs, gbd, cap[5];
};
KH_HD inline MosAC mos_ac(const Lane& L, const int32_t* e) {
  MosAC a;
  double vgs = L.S(e, MS_AVGS), vds = L.S(e, MS_AVDS), vbs = L.S(e, MS_AVBS);
  double vgd = L.S(e, MS_AVGD), vbd = L.S(e, MS_AVBD);
  MosIds ch = mos_channel(L, e, vgs, vds, vbs);
  if (ch.md > 0) {
    a.gch[0] = ch.gm; a.gch[1] = ch.gds; a.gch[2] = ch.gmbs;
  } else {
    a.gch[0] = -ch.gm; a.gch[1] = ch.gm + ch.gds + ch.gmbs; a.gch[2] = -ch.gmbs;
  }
  double gmin = L.S(e, MS_GMIN), tmp;
  mos_bulk_diode(L, e, vbs, L.P(e, 10), gmin, &tmp, &a.gbs);
  mos_bulk_diode(L, e, vbd, L.P(e, 9), gmin, &tmp, &a.gbd);
  double mc[3];
  mos_meyer_caps(L, e, vgs, vgd, ch.von, ch.vdsat, ch.md, mc);
  double q, cbs, cbd;
  mos_bulk_charge(L, e, vbs, L.P(e, 15), L.P(e, 16), &q, &cbs);
  mos_bulk_charge(L, e, vbd, L.P(e, 13), L.P(e, 14), &q, &cbd);
  a.cap[0] = mc[0] + L.P(e, 27);
  a.cap[1] = mc[1] + L.P(e, 28);
  a.cap[2] = mc[2] + L.P(e, 29);
  a.cap[3] = cbs;
  a.cap[4] = cbd;
  return a;
}

// BJT small-signal admittances at the accepted point (device/bjt.go:690-705).
struct BjtAC {
  double itBE, itBC, gpi, gmu, gx;
  kcplx ybe, ybeBC, ybc, ybx, ycs;
};
KH_HD inline BjtAC bjt_ac(const Lane& L, const int32_t* e, double omega) {
  double v[4] = {L.S(e, QS_ACC + 0), L.S(e, QS_ACC + 1), L.S(e, QS_ACC + 2), L.S(e, QS_ACC + 3)};
  double gmin = L.S(e, QS_GMIN);
  BjtDC d = bjt_eval_dc(L, e, v[0], v[1], gmin);
  BjtQ ch = bjt_charges(L, e, v, d);
  double isub, gsub;
  bjt_junction(L.P(e, QP_ISS), v[3], L.P(e, QP_NS) * L.P(e, QP_VT), &isub, &gsub);
  gsub += gmin;
  BjtAC a;
  a.itBE = d.itBE; a.itBC = d.itBC; a.gpi = d.gpi; a.gmu = d.gmu; a.gx = d.gx;
  a.ybe = {0, omega * ch.c[0]};
  a.ybeBC = {0, omega * ch.dqbeDvbc};
  a.ybc = {0, omega * ch.c[1]};
  a.ybx = {0, omega * ch.c[2]};
  a.ycs = {gsub, omega * ch.c[3]};
  return a;
}
KH_HD inline bool cnonzero(kcplx y) { return y.re != 0 || y.im != 0; }
// acBranch, device/bjt.go:718-731 (terminal indices)
KH_HD inline void bjt_ac_branch(ACLane& A, const int32_t* e, int ta, int tb, kcplx y1, int p1, int n1, kcplx y2,
                                int p2, int n2) {
  if (cnonzero(y1)) {
    A.add(e, ta, p1, y1); A.add(e, ta, n1, -y1); A.add(e, tb, p1, -y1); A.add(e, tb, n1, y1);
  }
  if (cnonzero(y2)) {
    A.add(e, ta, p2, y2); A.add(e, ta, n2, -y2); A.add(e, tb, p2, -y2); A.add(e, tb, n2, y2);
  }
}

// LoadAC of every element (device/passive.go, reactive.go:355-401,
// source.go:186-217, diode.go:409-420, mosfet.go:779-814, bjt.go:690-716), in
// element order.
KH_HD inline void ac_load(ACLane& A, int nel) {
  const Lane& L = A.dev;
  for (int k = 0; k < nel; k++) {
    const int32_t* e = L.et + (int64_t)k * ET_W;
    switch (e[ET_KIND]) {
      case K_R: A.stampY(e, 0, 1, {L.P(e, 1), 0}); break;
      case K_C: A.stampY(e, 0, 1, {0, A.omega * L.P(e, 0)}); break;
      case K_L:
        A.add(e, 0, 2, {1, 0});
        A.add(e, 1, 2, {-1, 0});
        A.add(e, 2, 0, {1, 0});
        A.add(e, 2, 1, {-1, 0});
        A.add(e, 2, 2, {0, -A.omega * L.P(e, 0)});
        break;
      case K_K: {  // Mutual.LoadAC, device/mutual.go:141-148: -jwM between the branch rows
        const double im = -A.omega * L.P(e, 0);
        A.add(e, 0, 1, {-0.0, im});  // complex(-0, -wM) - complex(0, 0)
        A.add(e, 1, 0, {0.0, im});   // complex(-0, -wM) + complex(0, 0)
        break;
      }
      case K_V:
        A.add(e, 0, 2, {1, 0});
        A.add(e, 1, 2, {-1, 0});
        A.add(e, 2, 0, {1, 0});
        A.add(e, 2, 1, {-1, 0});
        A.add_rhs(L.T(e, 2), source_ac(L, e));
        break;
      case K_I: {
        kcplx s = source_ac(L, e);
        A.add_rhs(L.T(e, 0), -s);
        A.add_rhs(L.T(e, 1), s);
        break;
      }
      case K_D: {  // terms: 0 a, 1 k, 2 ai; state 10: the accepted junction voltage
        double v = L.S(e, 10), id, g0, q, cc;
        diode_current(L, e, v, &id, &g0);
        diode_charge(L, e, v, id, g0, &q, &cc);
        A.stampY(e, 2, 1, {g0 + L.S(e, 7), A.omega * cc});  // state 7: the gmin of the last load
        if (L.P(e, 14) > 0) A.stampY(e, 0, 2, {L.P(e, 14), 0});
        break;
      }
      case K_M: {  // terms: 0 d, 1 s, 2 g, 3 b, 4 dp, 5 sp
        MosAC m = mos_ac(L, e);
        const int ctrlP[3] = {2, 4, 3};
        for (int j = 0; j < 3; j++) {  // acBranch (dp -> sp) on (ctrl - sp), skipped when zero
          if (m.gch[j] == 0) continue;
          kcplx y = {m.gch[j], 0};
          A.add(e, 4, ctrlP[j], y);
          A.add(e, 4, 5, -y);
          A.add(e, 5, ctrlP[j], -y);
          A.add(e, 5, 5, y);
        }
        A.stampY(e, 3, 5, {m.gbs, 0});
        A.stampY(e, 3, 4, {m.gbd, 0});
        const int ends[5][2] = {{2, 5}, {2, 4}, {2, 3}, {3, 5}, {3, 4}};
        for (int j = 0; j < 5; j++) A.stampY(e, ends[j][0], ends[j][1], {0, A.omega * m.cap[j]});
        if (L.T(e, 4) != L.T(e, 0)) A.stampY(e, 0, 4, {L.P(e, 7), 0});
        if (L.T(e, 5) != L.T(e, 1)) A.stampY(e, 1, 5, {L.P(e, 8), 0});
        break;
      }
      case K_Q: {  // terms: 0 c, 1 e, 2 b, 3 s, 4 cp, 5 bp, 6 ep (device/bjt.go:690-716)
        BjtAC a = bjt_ac(L, e, A.omega);
        const int sa = bjt_sa(L, e), sb = bjt_sb(L, e);
        const kcplx z = {0, 0};
        bjt_ac_branch(A, e, 4, 6, {a.itBE, 0}, 5, 6, {a.itBC, 0}, 5, 4);
        bjt_ac_branch(A, e, 5, 6, kcplx{a.gpi, 0} + a.ybe, 5, 6, a.ybeBC, 5, 4);
        bjt_ac_branch(A, e, 5, 4, kcplx{a.gmu, 0} + a.ybc, 5, 4, z, 0, 0);
        bjt_ac_branch(A, e, 2, 4, a.ybx, 2, 4, z, 0, 0);
        bjt_ac_branch(A, e, sa, sb, a.ycs, sa, sb, z, 0, 0);
        if (L.T(e, 5) != L.T(e, 2)) A.stampY(e, 2, 5, {a.gx, 0});
        if (L.T(e, 4) != L.T(e, 0)) A.stampY(e, 0, 4, {L.P(e, QP_GCPR), 0});
        if (L.T(e, 6) != L.T(e, 1)) A.stampY(e, 1, 6, {L.P(e, QP_GEPR), 0});
        break;
      }
      case K_E:
      case K_F:
      case K_G:
      case K_H: {  // the DC stamps with a real gain (device/controlled.go:138-143)
        const double g = L.P(e, 0);
        const int kind = e[ET_KIND];
        if (kind == K_E || kind == K_H) {
          A.add(e, 0, 2, {1, 0});
          A.add(e, 1, 2, {-1, 0});
          A.add(e, 2, 0, {1, 0});
          A.add(e, 2, 1, {-1, 0});
          A.add(e, 2, 3, {-g, 0});
          if (kind == K_E) A.add(e, 2, 4, {g, 0});
        } else if (kind == K_G) {
          A.add(e, 0, 3, {g, 0});
          A.add(e, 0, 4, {-g, 0});
          A.add(e, 1, 3, {-g, 0});
          A.add(e, 1, 4, {g, 0});
        } else {
          A.add(e, 0, 3, {g, 0});
          A.add(e, 1, 3, {-g, 0});
        }
        break;
      }
    }
  }
}

// Fields of one frequency's average-power ledger (output.ACLedger), in this order.
enum { AL_INFLOW = 0, AL_OUTFLOW, AL_DISSIPATION, AL_REACTIVE, AL_MODEL_SUPPLIED, AL_RESIDUAL, AL_FIELDS };

// Complex power absorbed by every element, summed (sumS) and in magnitude
// (scaleS): OperatingAC + ACPower (analysis/ac.go:152-169). Optionally each
// element's current phasor (icur[k * istride]) and the frequency's ledger
// (led[AL_FIELDS], analysis/ac.go:160-173): the average power Re S drives each
// element's modeled exterior port (heat out of resistors and junction devices,
// the sources' power in from the exterior), Im S of the energy-storing
// elements sums into the reactive power, negative resistances supply Re S.
KH_HD inline void ac_power(const ACLane& A, int nel, const kcplx* x, kcplx* sumS, double* scaleS,
                           kcplx* icur = nullptr, int64_t istride = 1, double* led = nullptr) {
  const Lane& L = A.dev;
  kcplx sum = {0, 0};
  double scale = 0;
  double inflow = 0, outflow = 0, diss = 0, reactive = 0, sup = 0;
  for (int k = 0; k < nel; k++) {
    const int32_t* e = L.et + (int64_t)k * ET_W;
    kcplx s;
    if (e[ET_KIND] == K_M) {  // ACPower, device/mosfet.go:817-846
      MosAC m = mos_ac(L, e);
      kcplx gg = A.xat(x, L.T(e, 2)), bb = A.xat(x, L.T(e, 3)), dd = A.xat(x, L.T(e, 4)),
            ss = A.xat(x, L.T(e, 5));
      kcplx vgs = gg - ss, vds = dd - ss, vbs = bb - ss;
      kcplx ich = kcplx{m.gch[0], 0} * vgs + kcplx{m.gch[1], 0} * vds + kcplx{m.gch[2], 0} * vbs;
      kcplx ibs = kcplx{m.gbs, 0} * vbs;
      kcplx ibd = kcplx{m.gbd, 0} * (bb - dd);
      kcplx jw = {0, A.omega};
      kcplx iqs[5] = {jw * kcplx{m.cap[0], 0} * vgs, jw * kcplx{m.cap[1], 0} * (gg - dd),
                      jw * kcplx{m.cap[2], 0} * (gg - bb), jw * kcplx{m.cap[3], 0} * vbs,
                      jw * kcplx{m.cap[4], 0} * (bb - dd)};
      kcplx id = ich - ibd - iqs[1] - iqs[4];
      kcplx is = -ich - ibs - iqs[0] - iqs[3];
      kcplx ig = iqs[0] + iqs[1] + iqs[2];
      kcplx ib = ibs + ibd - iqs[2] + iqs[3] + iqs[4];
      s = A.xat(x, L.T(e, 0)) * conj(id) + gg * conj(ig) + A.xat(x, L.T(e, 1)) * conj(is) + bb * conj(ib);
      s = {s.re / 2.0, s.im / 2.0};
      if (icur) icur[(int64_t)k * istride] = id;
    } else if (e[ET_KIND] == K_Q) {  // acCurrents + ACPower, device/bjt.go:734-765
      BjtAC a = bjt_ac(L, e, A.omega);
      kcplx xb = A.xat(x, L.T(e, 2)), xcp = A.xat(x, L.T(e, 4)), xbp = A.xat(x, L.T(e, 5)), xep = A.xat(x, L.T(e, 6));
      kcplx vbe = xbp - xep, vbc = xbp - xcp, vbx = xb - xcp;
      kcplx vcs = A.xat(x, L.T(e, bjt_sa(L, e))) - A.xat(x, L.T(e, bjt_sb(L, e)));
      kcplx it = kcplx{a.itBE, 0} * vbe + kcplx{a.itBC, 0} * vbc;
      kcplx ibe = (kcplx{a.gpi, 0} + a.ybe) * vbe + a.ybeBC * vbc;
      kcplx ibc = (kcplx{a.gmu, 0} + a.ybc) * vbc;
      kcplx ix = a.ybx * vbx, isb = a.ycs * vcs;
      kcplx ic, ib, ie = -it - ibe, is;
      if (L.P(e, QP_SUBS) > 0) {
        ic = it - ibc - ix - isb; ib = ibe + ibc + ix; is = isb;
      } else {
        ic = it - ibc - ix; ib = ibe + ibc + ix + isb; is = -isb;
      }
      s = A.xat(x, L.T(e, 0)) * conj(ic) + xb * conj(ib) + A.xat(x, L.T(e, 1)) * conj(ie) +
          A.xat(x, L.T(e, 3)) * conj(is);
      s = {s.re / 2.0, s.im / 2.0};
      if (icur) icur[(int64_t)k * istride] = ic;
    } else {
      // a K element has no phasors of its own (device/mutual.go:150-157)
      kcplx v = e[ET_KIND] == K_K ? kcplx{0, 0} : A.xat(x, L.T(e, 0)) - A.xat(x, L.T(e, 1));
      kcplx i = {0, 0};
      switch (e[ET_KIND]) {
        case K_R: i = v * kcplx{L.P(e, 1), 0}; break;
        case K_C: i = v * kcplx{0, A.omega * L.P(e, 0)}; break;
        case K_L:
        case K_V:
        case K_E:
        case K_H: i = A.xat(x, L.T(e, 2)); break;
        case K_I: i = source_ac(L, e); break;
        case K_G: i = kcplx{L.P(e, 0), 0} * (A.xat(x, L.T(e, 3)) - A.xat(x, L.T(e, 4))); break;  // OperatingAC
        case K_F: i = kcplx{L.P(e, 0), 0} * A.xat(x, L.T(e, 3)); break;
        case K_D: {
          kcplx vj = A.xat(x, L.T(e, 2)) - A.xat(x, L.T(e, 1));
          double vd = L.S(e, 10), id, g0, q, cc;
          diode_current(L, e, vd, &id, &g0);
          diode_charge(L, e, vd, id, g0, &q, &cc);
          i = vj * kcplx{g0 + L.S(e, 7), A.omega * cc};
          break;
        }
      }
      s = kcplx{0.5 * v.re, 0.5 * v.im} * conj(i);  // 0.5*V times conj(I), in that order
      if (icur) icur[(int64_t)k * istride] = i;
    }
    sum = sum + s;
    scale += go_hypot(s.re, s.im);
    if (led) {
      const int kind = e[ET_KIND];
      const double pw = s.re;
      if (kind == K_C || kind == K_L || kind == K_D || kind == K_M || kind == K_Q || kind == K_K) reactive += s.im;
      double port = 0;
      int ch = -1;
      if ((kind == K_R && L.P(e, 0) > 0) || kind == K_D || kind == K_M || kind == K_Q) {
        port = -pw;
        ch = 2;
      } else if (kind == K_V || kind == K_I) {
        port = -pw;
        ch = 0;
      }
      if (ch >= 0 && port != 0) {  // Accumulator.AddPort
        if (ch == 2 && port < 0) diss += -port;
        else if (port > 0) inflow += port;
        else outflow += -port;
      }
      if (((kind == K_R && L.P(e, 0) < 0) || is_controlled(kind)) && pw != 0) sup += -pw;  // AddModelSupplied
    }
  }
  *sumS = sum;
  *scaleS = scale;
  if (led) {
    led[AL_INFLOW] = inflow;
    led[AL_OUTFLOW] = outflow;
    led[AL_DISSIPATION] = diss;
    led[AL_REACTIVE] = reactive;
    led[AL_MODEL_SUPPLIED] = sup;
    led[AL_RESIDUAL] = inflow + sup - outflow - diss;
  }
}

}  // namespace kh
