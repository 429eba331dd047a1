// kh_mos23.h — MOSFET level 2 and 3 drain currents for one lane
// (device/mosfet_level23.go, device/dual.go).
//
// The reference writes both currents once in forward-mode dual numbers over
// (vgs, vds, vbs), so gm, gds and gmbs are the exact derivatives of the current
// that is stamped; KDual restates that type operation for operation, and the
// two level functions restate the Go functions expression for expression
// (same association, same special cases), with the device math library in
// place of Go's math package.
#pragma once
#include "kh_lu.h"

namespace kh {

constexpr double kEpsSi = 1.035943139907e-10;  // epsSi = 11.7 * 8.854214871e-12, folded exactly

// Forward-mode dual number over (vgs, vds, vbs), device/dual.go:17-106: the
// MOSFET level 2 and 3 currents are evaluated in it, so gm, gds and gmbs are
// their exact derivatives. Each operation restates the Go method's arithmetic.
struct KDual {
  double v = 0, d[3] = {0, 0, 0};
};
KH_HD inline KDual dcst(double v) { KDual r; r.v = v; return r; }
KH_HD inline KDual dvar(double v, int k) { KDual r; r.v = v; r.d[k] = 1; return r; }
KH_HD inline KDual operator+(const KDual& a, const KDual& b) {
  KDual r; r.v = a.v + b.v;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
KH_HD inline KDual operator-(const KDual& a, const KDual& b) {
  KDual r; r.v = a.v - b.v;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
KH_HD inline KDual addc(KDual a, double c) { a.v += c; return a; }
KH_HD inline KDual dneg(const KDual& a) {
  KDual r; r.v = -a.v;
  for (int k = 0; k < 3; k++) r.d[k] = -a.d[k];
  return r;
}
KH_HD inline KDual scale(const KDual& a, double c) {
  KDual r; r.v = a.v * c;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] * c;
  return r;
}
KH_HD inline KDual operator*(const KDual& a, const KDual& b) {
  KDual r; r.v = a.v * b.v;
  for (int k = 0; k < 3; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
KH_HD inline KDual operator/(const KDual& a, const KDual& b) {
  KDual r; double q = a.v / b.v; r.v = q;
  for (int k = 0; k < 3; k++) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
  return r;
}
KH_HD inline KDual chain(const KDual& a, double f, double df) {
  KDual r; r.v = f;
  for (int k = 0; k < 3; k++) r.d[k] = df * a.d[k];
  return r;
}
KH_HD inline KDual dsqrt(const KDual& a) { double s = sqrt(a.v); return chain(a, s, 0.5 / s); }
KH_HD inline KDual dexp(const KDual& a) { double e = exp(a.v); return chain(a, e, e); }
KH_HD inline KDual dlog(const KDual& a) { return chain(a, log(a.v), 1 / a.v); }
KH_HD inline KDual dcos(const KDual& a) { return chain(a, cos(a.v), -sin(a.v)); }
KH_HD inline KDual datan(const KDual& a) { return chain(a, atan(a.v), 1 / (1 + a.v * a.v)); }
KH_HD inline KDual dabs(const KDual& a) { return a.v < 0 ? dneg(a) : a; }
KH_HD inline KDual powc(const KDual& a, double c) { double p = pow(a.v, c); return chain(a, p, c * p / a.v); }
KH_HD inline KDual dmax(const KDual& a, const KDual& b) { return b.v > a.v ? b : a; }
KH_HD inline KDual dmin(const KDual& a, const KDual& b) { return b.v < a.v ? b : a; }

// MOSFET level 2 and 3 parameters and derived constants (the M layout's
// level block, device/mosfet.go:128-324 and device/mosfet_levels.go:22-55).
struct MosLevelP {
  int level = 1;
  double NFS = 0, XJ = 0, UCRIT = 1e4, UEXP = 0, VMAX = 0, NEFF = 1, DELTA = 0, THETA = 0, ETA = 0, KAPPA = 0, UO = 600;
  double leff = 0, weff = 0, cox = 0, xd = 0, alpha = 0, cfs = 0, factor = 0, narrow = 0, eta3 = 0, sqrtPB = 0;
  bool hasNSUB = false;
};

struct IdsOut {
  double cd = 0, gm = 0, gds = 0, gmbs = 0, von = 0, vdsat = 0;
};
KH_HD inline IdsOut fromKDual(const KDual& id, const KDual& von, const KDual& vdsat) {
  IdsOut o;
  o.cd = id.v; o.gm = id.d[0]; o.gds = id.d[1]; o.gmbs = id.d[2]; o.von = von.v; o.vdsat = vdsat.v;
  return o;
}

// cbrtAbs, device/mosfet_level23.go:263-270
KH_HD inline KDual cbrtAbs(KDual a) {
  a = dabs(a);
  if (a.v == 0) return KDual();
  return powc(a, 1.0 / 3);
}

// level2VelocitySat, device/mosfet_level23.go:195-261: the Baum–Beneking
// quartic for x = sqrt(vdsat + phi - vbs), Ferrari's closed form; the smallest
// positive root with |poly| <= 1e-6.
KH_HD inline bool level2VelocitySat(const KDual& gammad, const KDual& v1, const KDual& v2, const KDual& invxv, const KDual& sarg3,
                              KDual* xout) {
  KDual xv = dcst(1) / invxv;
  KDual a1 = scale(gammad, 1 / 0.75);
  KDual b1 = scale(v1 + xv, -2);
  KDual c1 = scale(gammad * xv, -2);
  KDual d1 = scale(v1 * (v2 + xv), 2) - v2 * v2 - scale(gammad * sarg3, 4.0 / 3.0);
  KDual a = dneg(b1);
  KDual b = a1 * c1 - scale(d1, 4);
  KDual c = dneg(d1) * (a1 * a1 - scale(b1, 4)) - c1 * c1;
  KDual r = scale(a * a, -1.0 / 3) + b;
  KDual s = scale(a * a * a, 2.0 / 27) - scale(a * b, 1.0 / 3) + c;
  KDual r3 = r * r * r;
  KDual s2 = s * s;
  KDual pp = scale(s2, 0.25) + scale(r3, 1.0 / 27);
  KDual p0 = dabs(pp);
  KDual p2 = dsqrt(p0);
  KDual y3;
  if (pp.v < 0) {
    KDual ro = dexp(scale(dlog(dsqrt(scale(s2, 0.25) + p0)), 1.0 / 3));
    KDual fi = datan(scale(p2, -2) / s);
    y3 = scale(ro * dcos(scale(fi, 1.0 / 3)), 2) - scale(a, 1.0 / 3);
  } else {
    y3 = cbrtAbs(scale(s, -0.5) + p2) + cbrtAbs(scale(s, -0.5) - p2) - scale(a, 1.0 / 3);
  }
  KDual ra = scale(a1 * a1, 0.25) - b1 + y3;
  KDual rb = scale(y3 * y3, 0.25) - d1;
  if (!(ra.v >= 0) || !(rb.v >= 0)) return false;
  KDual a3 = dsqrt(ra), b3 = dsqrt(rb);
  const double sig1[4] = {1, -1, 1, -1}, sig2[4] = {1, 1, -1, -1};
  bool ok = false;
  KDual x;
  for (int i = 0; i < 4; i++) {
    KDual a4 = scale(a1, 0.5) + scale(a3, sig1[i]);
    KDual b4 = scale(y3, 0.5) + scale(b3, sig2[i]);
    KDual delta4 = scale(a4 * a4, 0.25) - b4;
    if (!(delta4.v >= 0)) continue;
    KDual t = dsqrt(delta4);
    KDual roots[2] = {scale(a4, -0.5) + t, scale(a4, -0.5) - t};
    for (const KDual& rt : roots) {
      if (!(rt.v > 0)) continue;
      double x2 = rt.v * rt.v;
      double poly = x2 * x2 + a1.v * x2 * rt.v + b1.v * x2 + c1.v * rt.v + d1.v;
      if (!(fabs(poly) <= 1e-6)) continue;
      if (!ok || rt.v < x.v) { x = rt; ok = true; }
    }
  }
  *xout = x;
  return ok;
}

// idsLevel2, device/mosfet_level23.go:31-193 (SPICE2 level 2, Grove–Frohman).
KH_HD inline IdsOut idsLevel2(const MosLevelP& lv, double vbi, double beta, double vt, double PHI, double GAMMA,
                        double LAMBDA, double vgsV, double vdsV, double vbsV) {
  KDual G = dvar(vgsV, 0), D = dvar(vdsV, 1), B = dvar(vbsV, 2);
  double phi = PHI;
  double sphi = sqrt(phi);
  double sphi3 = phi * sphi;
  KDual phiMinVbs = addc(dneg(B), phi);
  KDual sarg, dsrgdb;
  if (B.v <= 0) {
    sarg = dsqrt(phiMinVbs);
    dsrgdb = dcst(-0.5) / sarg;
  } else {
    sarg = dcst(sphi) / addc(scale(B, 0.5 / phi), 1);
    dsrgdb = scale(sarg * sarg, -0.5 / sphi3);
  }
  KDual barg, dbrgdb;
  if (B.v - D.v <= 0) {
    barg = dsqrt(phiMinVbs + D);
    dbrgdb = dcst(-0.5) / barg;
  } else {
    barg = dcst(sphi) / addc(scale(B - D, 0.5 / phi), 1);
    dbrgdb = scale(barg * barg, -0.5 / sphi3);
  }
  double factor = lv.factor;
  double eta = 1 + factor;
  KDual vbin = addc(scale(phiMinVbs, factor), vbi);
  KDual gamasd = dcst(GAMMA), dgddvb = dcst(0);
  if (GAMMA > 0 || lv.hasNSUB) {
    if (lv.XJ > 0) {
      KDual xws = scale(sarg, lv.xd), xwd = scale(barg, lv.xd);
      KDual args = dsqrt(addc(scale(xws, 2 / lv.XJ), 1));
      KDual argd = dsqrt(addc(scale(xwd, 2 / lv.XJ), 1));
      double t = 0.5 * lv.XJ / lv.leff;
      KDual argss = scale(addc(args, -1), t), argsd = scale(addc(argd, -1), t);
      gamasd = scale(addc(dneg(argss + argsd), 1), GAMMA);
      KDual dbargs = scale(scale(dsrgdb, lv.xd) / args, 0.5 / lv.leff);
      KDual dbargd = scale(scale(dbrgdb, lv.xd) / argd, 0.5 / lv.leff);
      dgddvb = scale(dbargs + dbargd, -GAMMA);
    }
  }
  KDual vonD = vbin + gamasd * sarg;
  bool nfs = lv.NFS != 0;
  KDual xn, argg;
  if (nfs) {
    KDual cdonco = addc(dneg(gamasd * dsrgdb + dgddvb * sarg), factor);
    xn = addc(cdonco, 1 + lv.cfs);
    KDual vtxn = scale(xn, vt);
    vonD = vonD + vtxn;
    argg = dcst(1) / vtxn;
  } else if (G.v <= vbin.v) {
    IdsOut o;
    o.von = vonD.v;
    return o;  // cutoff
  }
  KDual vgst = G - vonD;
  KDual ufact = dcst(1);
  double crit = lv.UCRIT * 100 * kEpsSi / lv.cox;
  if (vgst.v > crit) ufact = powc(dcst(crit) / vgst, lv.UEXP);
  KDual ueff = scale(ufact, lv.UO * 1e-4);
  KDual vgsx = G;
  if (nfs) vgsx = dmax(G, vonD);
  KDual gammad = scale(gamasd, 1 / eta);
  KDual vgb = scale(vgsx - vbin, 1 / eta);
  KDual vdsatD = dcst(0);
  if (gammad.v > 0) {
    KDual argv = vgb + phiMinVbs;
    if (argv.v > 0) {
      KDual gammd2 = gammad * gammad;
      KDual arg1 = dsqrt(addc(scale(argv, 4) / gammd2, 1));
      vdsatD = dmax(vgb + scale(gammd2 * addc(dneg(arg1), 1), 0.5), dcst(0));
    }
  } else {
    vdsatD = dmax(vgb, dcst(0));
  }
  KDual sarg3 = sarg * sarg * sarg;
  if (lv.VMAX > 0) {
    KDual x;
    if (level2VelocitySat(gammad, vgb + phiMinVbs, phiMinVbs, scale(ueff, 1 / (lv.VMAX * lv.leff)), sarg3, &x))
      vdsatD = x * x - phiMinVbs;
  }
  KDual xlamda = dcst(LAMBDA);
  KDual bsarg, bodys;
  if (D.v != 0) {
    if (B.v - vdsatD.v <= 0) bsarg = dsqrt(vdsatD + phiMinVbs);
    else bsarg = dcst(sphi) / addc(scale(B - vdsatD, 0.5 / phi), 1);
    bodys = bsarg * bsarg * bsarg - sarg3;
    if (lv.hasNSUB && !(LAMBDA > 0)) {
      if (lv.VMAX <= 0) {
        KDual argv = scale(D - vdsatD, 0.25);
        KDual sargv = dsqrt(addc(argv * argv, 1));
        xlamda = dsqrt(argv + sargv) * (dcst(lv.xd) / scale(D, lv.leff));
      } else {
        double xdv = lv.xd / sqrt(lv.NEFF);
        KDual xlv = dcst(lv.VMAX * xdv * 0.5) / ueff;
        KDual argv = dmax(D - vdsatD, dcst(0));
        KDual xls = dsqrt(xlv * xlv + argv);
        xlamda = (xls - xlv) * (dcst(xdv) / scale(D, lv.leff));
      }
    }
  }
  double xwb = lv.xd * lv.sqrtPB;
  double xld = lv.leff - xwb;
  KDual clfact = addc(dneg(xlamda * D), 1);
  if (!lv.hasNSUB) xwb = 0.25e-6;
  KDual xleff = scale(clfact, lv.leff);
  if (xleff.v < xwb) {
    KDual deltal = scale(xlamda * D, lv.leff);
    xleff = dcst(xwb) / addc(scale(addc(deltal, -xld), 1 / xwb), 1);
    clfact = scale(xleff, 1 / lv.leff);
  }
  KDual beta1 = scale(ufact, beta) / clfact;
  if (D.v <= 1e-10) {
    KDual g;
    if (G.v > vonD.v) g = beta1 * (G - vbin - gamasd * sarg);
    else if (nfs) g = beta1 * (vonD - vbin - gamasd * sarg) * dexp(argg * (G - vonD));
    IdsOut o;
    o.gds = g.v; o.von = vonD.v; o.vdsat = vdsatD.v;
    return o;
  }
  KDual id;
  if (nfs && G.v <= vonD.v) {
    if (vdsatD.v <= 0) {
      IdsOut o;
      o.von = vonD.v; o.vdsat = vdsatD.v;
      return o;
    }
    KDual vdson = dmin(vdsatD, D);
    KDual body = barg * barg * barg - sarg3;
    if (D.v > vdsatD.v) body = bodys;
    KDual cdson = beta1 * ((vonD - vbin - scale(vdson, 0.5 * eta)) * vdson - scale(gamasd * body, 1 / 1.5));
    id = cdson * dexp(argg * (G - vonD));
  } else if (D.v <= vdsatD.v) {
    KDual body = barg * barg * barg - sarg3;
    id = beta1 * ((G - vbin - scale(D, 0.5 * eta)) * D - scale(gamasd * body, 1 / 1.5));
  } else {
    id = beta1 * ((G - vbin - scale(vdsatD, 0.5 * eta)) * vdsatD - scale(gamasd * bodys, 1 / 1.5));
  }
  return fromKDual(id, vonD, vdsatD);
}

// idsLevel3, device/mosfet_level23.go:282-378 (SPICE2 level 3, semi-empirical).
KH_HD inline IdsOut idsLevel3(const MosLevelP& lv, double vbi, double beta, double vt, double PHI, double GAMMA,
                        double LD, double vgsV, double vdsV, double vbsV) {
  const double coeff0 = 0.0631353, coeff1 = 0.8013292, coeff2 = -0.01110777;
  KDual G = dvar(vgsV, 0), D = dvar(vdsV, 1), B = dvar(vbsV, 2);
  double phi = PHI;
  double leff = lv.leff, weff = lv.weff;
  KDual phibs, sqphbs;
  if (B.v <= 0) {
    phibs = addc(dneg(B), phi);
    sqphbs = dsqrt(phibs);
  } else {
    sqphbs = dcst(sqrt(phi)) / addc(scale(B, 1 / (2 * phi)), 1);
    phibs = sqphbs * sqphbs;
  }
  KDual fshort = dcst(1);
  if (lv.XJ != 0 && lv.xd != 0) {
    KDual wponxj = scale(sqphbs, lv.xd / lv.XJ);
    KDual wconxj = addc(scale(wponxj * wponxj, coeff2) + scale(wponxj, coeff1), coeff0);
    double djonxj = LD / lv.XJ;
    KDual arga = addc(wconxj, djonxj);
    KDual argc = wponxj / addc(wponxj, 1);
    KDual argb = dsqrt(addc(dneg(argc * argc), 1));
    fshort = addc(scale(addc(arga * argb, -djonxj), -lv.XJ / leff), 1);
  }
  KDual gammas = scale(fshort, GAMMA);
  KDual fbodys = scale(gammas / sqphbs, 0.25);
  KDual fbody = addc(fbodys, lv.narrow / weff);
  KDual qbonco = gammas * sqphbs + scale(phibs, lv.narrow / weff);
  KDual vth = addc(scale(D, -lv.eta3), vbi) + qbonco;
  KDual vonD = vth;
  bool nfs = lv.NFS != 0;
  KDual xn;
  if (nfs) {
    xn = addc(qbonco / scale(phibs, 2), 1 + lv.cfs);
    vonD = vth + scale(xn, vt);
  } else if (G.v <= vonD.v) {
    IdsOut o;
    o.von = vonD.v;
    return o;
  }
  KDual vgsx = dmax(G, vonD);
  KDual onfg = addc(scale(vgsx - vth, lv.THETA), 1);
  KDual fgate = dcst(1) / onfg;
  KDual us = scale(fgate, lv.UO * 1e-4);
  KDual onfbdy = dcst(1) / addc(fbody, 1);
  KDual arga = (vgsx - vth) * onfbdy;
  KDual vdsatD = arga;
  KDual vdsc;
  if (lv.VMAX > 0) {
    vdsc = dcst(leff * lv.VMAX) / us;
    KDual argb = dsqrt(arga * arga + vdsc * vdsc);
    vdsatD = arga + vdsc - argb;
  }
  KDual vdsx = dmin(D, vdsatD);
  KDual betaD = scale(fgate, beta);
  if (vdsx.v == 0) {
    KDual g = betaD * (vgsx - vth);
    if (nfs && G.v < vonD.v) g = g * dexp((G - vonD) / scale(xn, vt));
    IdsOut o;
    o.gds = g.v; o.von = vonD.v; o.vdsat = vdsatD.v;
    return o;
  }
  KDual cdo = vgsx - vth - scale(addc(fbody, 1) * vdsx, 0.5);
  KDual id = betaD * cdo * vdsx;
  KDual fdrain;
  if (lv.VMAX > 0) {
    fdrain = dcst(1) / addc(vdsx / vdsc, 1);
    id = id * fdrain;
  }
  bool clm = false;
  KDual delxl;
  if (D.v <= vdsatD.v) {
    if (lv.VMAX <= 0 && lv.alpha != 0) {
      KDual r = D / vdsatD;
      KDual r2 = r * r;
      delxl = dsqrt(scale(vdsatD, lv.KAPPA * lv.alpha / 8)) * (r2 * r2);
      clm = true;
    }
  } else if (lv.VMAX <= 0) {
    delxl = dsqrt(scale(D - vdsatD + scale(vdsatD, 1.0 / 8), lv.KAPPA * lv.alpha));
    clm = true;
  } else if (lv.alpha != 0) {
    KDual gdsat = dmax(id * addc(dneg(fdrain), 1) / vdsc, dcst(1e-12));
    KDual emax = scale(id / gdsat, lv.KAPPA / leff);
    KDual a = scale(emax, 0.5 * lv.alpha);
    KDual argb = dsqrt(a * a + scale(D - vdsatD, lv.KAPPA * lv.alpha));
    delxl = argb - a;
    clm = true;
  }
  if (clm) {
    if (delxl.v > 0.5 * leff) delxl = addc(dcst(-leff * leff / 4) / delxl, leff);
    id = id / addc(scale(delxl, -1 / leff), 1);
  }
  if (nfs && G.v < vonD.v) id = id * dexp((G - vonD) / scale(xn, vt));
  return fromKDual(id, vonD, vdsatD);
}

// The level block of the M parameter layout (include/kronhip.h, 34..57).
KH_HD inline MosLevelP mos_level_params(const double* p) {
  MosLevelP lv;
  lv.level = (int)p[34]; lv.NFS = p[35]; lv.XJ = p[36]; lv.UCRIT = p[37]; lv.UEXP = p[38]; lv.VMAX = p[39];
  lv.NEFF = p[40]; lv.DELTA = p[41]; lv.THETA = p[42]; lv.ETA = p[43]; lv.KAPPA = p[44]; lv.UO = p[45];
  lv.leff = p[46]; lv.weff = p[47]; lv.cox = p[48]; lv.xd = p[49]; lv.alpha = p[50]; lv.cfs = p[51];
  lv.factor = p[52]; lv.narrow = p[53]; lv.eta3 = p[54]; lv.sqrtPB = p[55]; lv.hasNSUB = p[56] != 0;
  return lv;
}

}  // namespace kh
