// The equation of state on the device: mixture moments, the caloric polynomials,
// and the inversions (rho, e) -> T and (p, T) -> rho.
#pragma once

#include "pga_common.hpp"

namespace pga {

enum { EOS_CPG = 0, EOS_TPG = 1, EOS_REAL = 2 };

struct EosState {
  double p, T, rhoinv, rho, e, cp, gamma, c, Z;
};

// the composition sums a point's state is made from: R, cp(T)'s ln T coefficients
// premultiplied by each species' gas constant (or the constant cp), and h298
template <int NS, int EOS>
struct MixMom;

template <int NS>
struct MixMom<NS, EOS_TPG> {
  double R, h298, cpcoeff[NT];
};

template <int NS>
struct MixMom<NS, EOS_CPG> {
  double R, h298, cp;
};

// the real gas: the thermally perfect moments and the cubic's sums that do not move with T,
// b, c (volume shift) and da, each over Y_n / MW_n (scaled by MWmix where they are used)
template <int NS>
struct MixMom<NS, EOS_REAL> {
  double R, h298, cpcoeff[NT], b, c, da;
};

// eos/kernels/mixing.mako:3-30: Y(n) gives the carried ns - 1; the last is what they leave
// The ln T polynomials of a tpg gas carry coefficients of 1e6 whose terms cancel to cp ~ 1e3:
// one rounding more or less per step moves h by 1e-11 of itself. The EOS is therefore evaluated
// with a fixed rounding sequence: the composition sums and the Horner steps as explicit fused
// multiply-adds, everything else one rounding per written operation, in the reference's order.
// The oracle evaluates the same sequence, so the two agree to the bit.
template <int NS, int EOS, class YF>
__device__ __forceinline__ void mixmom(const Gas<NS> &g, const YF &Y, MixMom<NS, EOS> &m) {
#pragma clang fp contract(off)
  m.R = 0.0;
  m.h298 = 0.0;
  if constexpr (EOS != EOS_CPG) {
#pragma unroll
    for (int t = 0; t < NT; t++) m.cpcoeff[t] = 0.0;
  } else {
    m.cp = 0.0;
  }
  if constexpr (EOS == EOS_REAL) m.b = m.c = m.da = 0.0;
  double Ylast = 1.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    double Yn;
    if (n < NS - 1) {
      Yn = Y(n);
      Ylast -= Yn;
    } else {
      Yn = Ylast;
    }
    m.R = fma(Yn, g.Rsp(n), m.R);
    if constexpr (EOS != EOS_CPG) {
#pragma unroll
      for (int t = 0; t < NT; t++) m.cpcoeff[t] = fma(Yn, g.cppolyR(n, t), m.cpcoeff[t]);
    } else {
      m.cp = fma(Yn, g.cp0(n), m.cp);
    }
    m.h298 = fma(Yn, g.h298R(n), m.h298);
    if constexpr (EOS == EOS_REAL) {
      m.b = fma(Yn, g.cubicbMW(n), m.b);
      m.c = fma(Yn, g.vshiftMW(n), m.c);
      m.da = fma(Yn, g.cubicdaMW(n), m.da);
    }
  }
}

// mixmom for two compositions at once (a face's two sides): each species' constants are loaded
// once and folded into both sums while they are live, rather than held (~80 values) across two
// separate passes; each side's sums are the single mixmom's to the bit
template <int NS, int EOS, class YA, class YB>
__device__ __forceinline__ void mixmom2(const Gas<NS> &g, const YA &Ya, const YB &Yb, MixMom<NS, EOS> &ma,
                                        MixMom<NS, EOS> &mb) {
#pragma clang fp contract(off)
  ma.R = mb.R = 0.0;
  ma.h298 = mb.h298 = 0.0;
  if constexpr (EOS != EOS_CPG) {
#pragma unroll
    for (int t = 0; t < NT; t++) ma.cpcoeff[t] = mb.cpcoeff[t] = 0.0;
  } else {
    ma.cp = mb.cp = 0.0;
  }
  if constexpr (EOS == EOS_REAL) ma.b = ma.c = ma.da = mb.b = mb.c = mb.da = 0.0;
  double la = 1.0, lb = 1.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    double ya, yb;
    if (n < NS - 1) {
      ya = Ya(n);
      yb = Yb(n);
      la -= ya;
      lb -= yb;
    } else {
      ya = la;
      yb = lb;
    }
    const double Rn = g.Rsp(n);
    ma.R = fma(ya, Rn, ma.R);
    mb.R = fma(yb, Rn, mb.R);
    if constexpr (EOS != EOS_CPG) {
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const double cn = g.cppolyR(n, t);
        ma.cpcoeff[t] = fma(ya, cn, ma.cpcoeff[t]);
        mb.cpcoeff[t] = fma(yb, cn, mb.cpcoeff[t]);
      }
    } else {
      const double cn = g.cp0(n);
      ma.cp = fma(ya, cn, ma.cp);
      mb.cp = fma(yb, cn, mb.cp);
    }
    const double hn = g.h298R(n);
    ma.h298 = fma(ya, hn, ma.h298);
    mb.h298 = fma(yb, hn, mb.h298);
    if constexpr (EOS == EOS_REAL) {
      const double bn = g.cubicbMW(n), vn = g.vshiftMW(n), dn = g.cubicdaMW(n);
      ma.b = fma(ya, bn, ma.b);
      mb.b = fma(yb, bn, mb.b);
      ma.c = fma(ya, vn, ma.c);
      mb.c = fma(yb, vn, mb.c);
      ma.da = fma(ya, dn, ma.da);
      mb.da = fma(yb, dn, mb.da);
    }
  }
}

// caloric/tdepcp/kernels/caloric.mako:4-15: cp and h per mass at T, h's coefficients
// built in the same descent as hcoeff[m] = cpcoeff[m] - (m + 1) hcoeff[m + 1]
template <class M>
__device__ __forceinline__ void ideal_cph(const M &m, double T, double &cp, double &h) {
#pragma clang fp contract(off)
  const double u = elog(T);
  double cpv = 0.0, hv = 0.0, hc = 0.0;
#pragma unroll
  for (int t = NT - 1; t >= 0; t--) {
    cpv = fma(cpv, u, m.cpcoeff[t]);
    hc = fma(-(double)(t + 1), hc, m.cpcoeff[t]);
    hv = fma(hv, u, hc);
  }
  cp = cpv;
  h = fma(hv, T, m.h298);
}

// ---------------------------------------------------------------------------
// the cubic real gas (eos/realgas/kernels/eos.mako, newton.mako): logarithms by elog, as the
// rest of the EOS
// ---------------------------------------------------------------------------
struct Cubic {
  double MWmix, bm, cm, S1, S2, am, dam, Astar, Bstar, Cstar, invRuT, Z, z1, z2, invslope, dZdT, v, I, logratio, q, qd;
};

// Soave's |1 + m (1 - sqrt(T / Tc))| opened as |(1 + m) - sqrt(T) m / sqrt(Tc)|
template <int NS>
__device__ __forceinline__ double modulus(const Gas<NS> &g, int n, double sqrtT) {
  return fabs(fma(-sqrtT, g.cubicmTc(n), g.cubiconepm(n)));
}

template <int NS>
__device__ __forceinline__ void cubicmix(const MixMom<NS, EOS_REAL> &m, Cubic &c) {
#pragma clang fp contract(off)
  c.MWmix = Ru / m.R;
  c.bm = m.b * c.MWmix;
  c.cm = m.c * c.MWmix;
  c.S2 = m.da * c.MWmix;
}

// k_ij = 0: am = S1^2 and dam/dT = -S1 S2 / sqrt(T)
template <int NS, class YF>
__device__ __forceinline__ void attraction(const Gas<NS> &g, Cubic &c, const YF &Y, double T) {
#pragma clang fp contract(off)
  const double sqrtT = sqrt(T);
  double s1 = 0.0, Ylast = 1.0;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) {
    const double Yn = Y(n);
    Ylast -= Yn;
    s1 = fma(Yn, g.cubicsqrtacMW(n) * modulus(g, n, sqrtT), s1);
  }
  s1 = fma(Ylast, g.cubicsqrtacMW(NS - 1) * modulus(g, NS - 1, sqrtT), s1);
  c.S1 = s1 * c.MWmix;
  c.am = c.S1 * c.S1;
  c.dam = -c.S1 * c.S2 / sqrtT;
}

template <int NS>
__device__ __forceinline__ void dimless(const Gas<NS> &g, Cubic &c, double p, double T) {
#pragma clang fp contract(off)
  const double u = g.cubic(0), w = g.cubic(1);
  c.invRuT = 1.0 / (Ru * T);
  const double pRT = p * c.invRuT;
  c.Astar = c.am * pRT * c.invRuT;
  c.Bstar = c.bm * pRT;
  c.Cstar = c.cm * pRT;
  const double B = c.Bstar, B2 = B * B;
  c.z1 = c.Astar + w * B2 - u * B - u * B2;
  c.z2 = -(1.0 + B - u * B);
}

__device__ __forceinline__ void rootslope(Cubic &c) {
#pragma clang fp contract(off)
  c.invslope = 1.0 / (3.0 * c.Z * c.Z + 2.0 * c.Z * c.z2 + c.z1);
}

// The cubic's root (eos/realgas/kernels/eos.mako:46-83): Cardano where one root is real; where
// three are, the largest by Newton from its upper bound 2 sqrt(Q) - z2/3 (the cubic rises convex
// above it, so the steps fall monotonically onto it) and the other two from the deflated
// quadratic; of the roots above B*, the least Gibbs energy departure's. The reference takes the
// three by the trigonometric form: these are the same roots in IEEE operations only (with dcbrt),
// so the oracle and the device agree to the bit.
template <int NS>
__device__ __forceinline__ void zsolve(const Gas<NS> &g, Cubic &c, double p, double T) {
#pragma clang fp contract(off)
  dimless(g, c, p, T);
  const double w = g.cubic(1), sig = g.cubic(2), eps = g.cubic(3);
  const double B = c.Bstar, B2 = B * B, z1 = c.z1, z2 = c.z2;
  const double z0 = -(c.Astar * B + w * B2 * (1.0 + B));
  const double Q = (z2 * z2 - 3.0 * z1) * (1.0 / 9.0);
  const double R = (2.0 * z2 * z2 * z2 - 9.0 * z2 * z1 + 27.0 * z0) * (1.0 / 54.0);
  const double M = R * R - Q * Q * Q;
  const double z2o3 = z2 * (1.0 / 3.0);
  if (M > 0.0) {
    const double S = -copysign(dcbrt(fabs(R) + sqrt(M)), R);
    c.Z = S + Q / S - z2o3;
  } else {
    double r0 = 2.0 * sqrt(Q) - z2o3;
    for (int k = 0; k < 100; k++) {
      const double f = ((r0 + z2) * r0 + z1) * r0 + z0;
      const double df = (3.0 * r0 + 2.0 * z2) * r0 + z1;
      if (!(df > 0.0)) break;
      const double dr = f / df;
      r0 -= dr;
      if (!(fabs(dr) > 1e-15 * fabs(r0))) break;
    }
    // Z^2 + bq Z + cq: the other two roots, the larger first, by the cancellation-free form
    const double bq = z2 + r0, cq = z1 + bq * r0, disc = bq * bq - 4.0 * cq;
    double r1 = r0, r2 = r0;
    if (disc >= 0.0) {
      const double qq = -0.5 * (bq + copysign(sqrt(disc), bq));
      const double ra = qq, rb = qq != 0.0 ? cq / qq : 0.0;
      r1 = fmax(ra, rb);
      r2 = fmin(ra, rb);
    }
    double Z = r0;
    const int above = (r0 > B) + (r1 > B) + (r2 > B);
    if (above > 1) {
      const double AoBse = c.Astar / (B * (sig - eps));
      double gmin = 0.0;
      bool found = false;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double z = k == 0 ? r0 : k == 1 ? r1 : r2;
        if (z <= B) continue;
        const double gz = z - 1.0 - elog(z - B) + AoBse * elog((z + eps * B) / (z + sig * B));
        if (!found || gz < gmin) {
          gmin = gz;
          Z = z;
          found = true;
        }
      }
    }
    c.Z = Z;
  }
  rootslope(c);
}

template <int NS>
__device__ __forceinline__ double dZ(const Gas<NS> &g, const Cubic &c, double dA, double dB) {
#pragma clang fp contract(off)
  const double u = g.cubic(0), w = g.cubic(1);
  const double B = c.Bstar, Z = c.Z, Z2 = Z * Z;
  const double dz0 = -(B * dA + (c.Astar + (2.0 * B + 3.0 * B * B) * w) * dB);
  const double dz1 = dA + (2.0 * B * (w - u) - u) * dB;
  const double dz2 = -(1.0 - u) * dB;
  return -(dz2 * Z2 + dz1 * Z + dz0) * c.invslope;
}

template <int NS>
__device__ __forceinline__ void departures(const Gas<NS> &g, Cubic &c, double T, double &hdep, double &cpdep) {
#pragma clang fp contract(off)
  const double u = g.cubic(0), w = g.cubic(1), sig = g.cubic(2), eps = g.cubic(3);
  const double Tinv = Ru * c.invRuT;
  const double dlnamdT = c.dam / c.am;
  c.dZdT = dZ(g, c, -2.0 * c.Astar * Tinv * (1.0 - 0.5 * T * dlnamdT), -c.Bstar * Tinv);
  const double x = c.Z / c.Bstar, xm1 = x - 1.0;
  const double D = x * x + u * x + w;
  c.logratio = elog((x + eps) / (x + sig));
  const double invBmRu = 1.0 / (c.bm * Ru);
  const double hres = invBmRu / (sig - eps) * (c.am * Tinv - c.dam) * c.logratio;
  c.qd = c.dam * invBmRu;
  c.q = c.am * invBmRu * Tinv;
  const double num = D - c.qd * xm1;
  cpdep = hres * 0.5 * T * dlnamdT + num * num / (D * D - c.q * (2.0 * x + u) * xm1 * xm1) - 1.0;
  hdep = hres + (c.Z - 1.0) - c.Cstar;
}

template <int NS>
__device__ __forceinline__ void gamma_c(const Gas<NS> &g, const Cubic &c, double p, double T, double rho, double cp,
                                        double invZt, double &gamma, double &cs) {
#pragma clang fp contract(off)
  const double invp = 1.0 / p, Tinv = Ru * c.invRuT;
  const double dZtdT = c.dZdT + c.Cstar * Tinv;
  const double dZdp = dZ(g, c, c.Astar * invp, c.Bstar * invp);
  const double dZtdp = dZdp - c.Cstar * invp;
  const double drhodp = rho * invp * (1.0 - p * dZtdp * invZt);
  const double r = -Tinv * (1.0 + T * dZtdT * invZt);
  const double csq = 1.0 / (drhodp - (T / cp) * r * r);
  gamma = drhodp * csq;
  cs = sqrt(fabs(csq));
}

// (rho, e) -> state: cpg closed form (eos/cpg/kernels/eos.mako:5-15), tpg Newton on
// e(T) = h(T) - R T from the cell's last T, at least three steps and at most twelve,
// stopping when |dT| <= 1e5 eps T (eos/tpg/kernels/eos.mako:8-20, newton.mako:3-11); the real
// gas's Newton at the cell's volume (eos/realgas/kernels/eos.mako:157-189, newton.mako:3-15)
template <int NS, int EOS, class YF>
__device__ __forceinline__ void fromrhoE(const Gas<NS> &g, const MixMom<NS, EOS> &m, const YF &Y, double rho, double e,
                                         double Tguess, EosState &s) {
#pragma clang fp contract(off)
  constexpr double tol = 1e5 * 2.220446049250313e-16;
  if constexpr (EOS == EOS_CPG) {
    const double cv = m.cp - m.R;
    s.rho = rho;
    s.rhoinv = 1.0 / rho;
    s.e = e;
    s.T = (e - m.h298) / cv;
    s.p = rho * m.R * s.T;
    s.cp = m.cp;
    s.gamma = m.cp / cv;
    s.c = sqrt(s.gamma * m.R * s.T);
    s.Z = 1.0;
  } else if constexpr (EOS == EOS_REAL) {
    const double sig = g.cubic(2), eps = g.cubic(3);
    Cubic c;
    cubicmix(m, c);
    c.v = c.MWmix / rho + c.cm;
    c.I = elog((c.v + sig * c.bm) / (c.v + eps * c.bm)) / ((sig - eps) * c.bm);
    const double invMWmix = 1.0 / c.MWmix;
    double T = (Tguess < 1.0) ? 300.0 : Tguess, cpi = 0.0, hi = 0.0;
    for (int k = 0; k < 12; k++) {
      attraction(g, c, Y, T);
      ideal_cph(m, T, cpi, hi);
      const double app = 0.5 * (c.S2 * c.S2 - c.dam) / T;
      const double f = hi - m.R * T - (c.am - T * c.dam) * c.I * invMWmix - e;
      const double cv = cpi - m.R + T * app * c.I * invMWmix;
      const double dT = f / cv;
      T -= dT;
      if (k >= 2 && fabs(dT) <= tol * T) break;
    }
    attraction(g, c, Y, T);
    ideal_cph(m, T, cpi, hi);
    const double vs = c.v + sig * c.bm, ve = c.v + eps * c.bm;
    const double p = Ru * T / (c.v - c.bm) - c.am / (vs * ve);
    dimless(g, c, p, T);
    c.Z = p * c.invRuT * c.v;
    rootslope(c);
    double hdep, cpdep;
    departures(g, c, T, hdep, cpdep);
    s.rho = rho;
    s.rhoinv = 1.0 / rho;
    s.e = e;
    s.p = p;
    s.T = T;
    s.Z = c.Z;
    s.cp = cpi + m.R * cpdep;
    const double invZt = 1.0 / (c.Z - c.Cstar);
    gamma_c(g, c, p, T, rho, s.cp, invZt, s.gamma, s.c);
  } else {
    double T = (Tguess < 1.0) ? 300.0 : Tguess, cp = 0.0, h = 0.0;
    for (int k = 0; k < 12; k++) {
      ideal_cph(m, T, cp, h);
      const double dT = (e - (h - m.R * T)) / (cp - m.R);
      T += dT;
      if (k >= 2 && fabs(dT) <= tol * T) break;
    }
    s.rho = rho;
    s.rhoinv = 1.0 / rho;
    s.e = e;
    s.T = T;
    s.p = rho * m.R * T;
    s.cp = cp;
    s.gamma = cp / (cp - m.R);
    s.c = sqrt(s.gamma * m.R * T);
    s.Z = 1.0;
  }
}

// (p, T) -> state: cpg eos.mako:17-27, tpg eos.mako:22-33, real gas eos/realgas/kernels/eos.mako:137-155
template <int NS, int EOS, class YF>
__device__ __forceinline__ void frompT(const Gas<NS> &g, const MixMom<NS, EOS> &m, const YF &Y, double p, double T,
                                       EosState &s) {
#pragma clang fp contract(off)
  if constexpr (EOS == EOS_REAL) {
    Cubic c;
    cubicmix(m, c);
    attraction(g, c, Y, T);
    double cpi, hi;
    ideal_cph(m, T, cpi, hi);
    zsolve(g, c, p, T);
    double hdep, cpdep;
    departures(g, c, T, hdep, cpdep);
    s.p = p;
    s.T = T;
    s.Z = c.Z;
    s.cp = cpi + m.R * cpdep;
    const double Zt = c.Z - c.Cstar, invZt = 1.0 / Zt, porho = Zt * m.R * T;
    s.rho = p * c.invRuT * c.MWmix * invZt;
    s.rhoinv = 1.0 / s.rho;
    s.e = hi + m.R * T * hdep - porho;
    gamma_c(g, c, p, T, s.rho, s.cp, invZt, s.gamma, s.c);
    return;
  } else {
    s.p = p;
    s.T = T;
    s.rho = p / (m.R * T);
    s.rhoinv = 1.0 / s.rho;
    if constexpr (EOS == EOS_CPG) {
      const double cv = m.cp - m.R;
      s.e = cv * T + m.h298;
      s.cp = m.cp;
      s.gamma = m.cp / cv;
    } else {
      double cp, h;
      ideal_cph(m, T, cp, h);
      s.e = h - m.R * T;
      s.cp = cp;
      s.gamma = cp / (cp - m.R);
    }
    s.c = sqrt(s.gamma * m.R * T);
    s.Z = 1.0;
  }
}

template <int EOS>
__device__ __forceinline__ void store_state(const View &st, int64_t o, const EosState &s) {
  if constexpr (EOS == EOS_REAL) st.at(o, 8) = s.Z;
  st.at(o, 0) = s.p;
  st.at(o, 1) = s.T;
  st.at(o, 2) = s.rhoinv;
  st.at(o, 3) = s.rho;
  st.at(o, 4) = s.e;
  st.at(o, 5) = s.cp;
  st.at(o, 6) = s.gamma;
  st.at(o, 7) = s.c;
}

__device__ __forceinline__ void load_state(const View &st, int64_t o, EosState &s) {
  s.p = st.at(o, 0);
  s.T = st.at(o, 1);
  s.rhoinv = st.at(o, 2);
  s.rho = st.at(o, 3);
  s.e = st.at(o, 4);
  s.cp = st.at(o, 5);
  s.gamma = st.at(o, 6);
  s.c = st.at(o, 7);
  s.Z = 1.0;
}

// euler/kernels/state/fromQ.mako:5-19 for the cell at o: Q -> state, T from state as the start
template <int NS, int EOS>
__device__ __forceinline__ void cell_fromQ(const Gas<NS> &g, const View &Q, const View &st, int64_t o) {
#pragma clang fp contract(off)
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  const double q1 = Q.at(o, 1), q2 = Q.at(o, 2), q3 = Q.at(o, 3);
  const double e = (Q.at(o, 4) - 0.5 * (q1 * q1 + q2 * q2 + q3 * q3) * rhoinv) * rhoinv;
  double Y[NS > 1 ? NS - 1 : 1];
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Y[n] = Q.at(o, 5 + n) * rhoinv;
  MixMom<NS, EOS> m;
  mixmom<NS, EOS>(g, [&](int n) { return Y[n]; }, m);
  EosState s;
  fromrhoE<NS, EOS>(g, m, [&](int n) { return Y[n]; }, rho, e, st.at(o, 1), s);
  store_state<EOS>(st, o, s);
}

}  // namespace pga
