// The equation of state on the device: mixture moments, the caloric polynomials,
// and the inversions (rho, e) -> T and (p, T) -> rho.
#pragma once

#include "pga_common.hpp"

namespace pga {

enum { EOS_CPG = 0, EOS_TPG = 1 };

struct EosState {
  double p, T, rhoinv, rho, e, cp, gamma, c;
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
  if constexpr (EOS == EOS_TPG) {
#pragma unroll
    for (int t = 0; t < NT; t++) m.cpcoeff[t] = 0.0;
  } else {
    m.cp = 0.0;
  }
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
    if constexpr (EOS == EOS_TPG) {
#pragma unroll
      for (int t = 0; t < NT; t++) m.cpcoeff[t] = fma(Yn, g.cppolyR(n, t), m.cpcoeff[t]);
    } else {
      m.cp = fma(Yn, g.cp0(n), m.cp);
    }
    m.h298 = fma(Yn, g.h298R(n), m.h298);
  }
}

// caloric/tdepcp/kernels/caloric.mako:4-15: cp and h per mass at T, h's coefficients
// built in the same descent as hcoeff[m] = cpcoeff[m] - (m + 1) hcoeff[m + 1]
template <int NS>
__device__ __forceinline__ void ideal_cph(const MixMom<NS, EOS_TPG> &m, double T, double &cp, double &h) {
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

// (rho, e) -> state: cpg closed form (eos/cpg/kernels/eos.mako:5-15), tpg Newton on
// e(T) = h(T) - R T from the cell's last T, at least three steps and at most twelve,
// stopping when |dT| <= 1e5 eps T (eos/tpg/kernels/eos.mako:8-20, newton.mako:3-11)
template <int NS, int EOS>
__device__ __forceinline__ void fromrhoE(const MixMom<NS, EOS> &m, double rho, double e, double Tguess,
                                         EosState &s) {
#pragma clang fp contract(off)
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
  } else {
    double T = (Tguess < 1.0) ? 300.0 : Tguess, cp = 0.0, h = 0.0;
    constexpr double tol = 1e5 * 2.220446049250313e-16;
    for (int k = 0; k < 12; k++) {
      ideal_cph<NS>(m, T, cp, h);
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
  }
}

// (p, T) -> state: cpg eos.mako:17-27, tpg eos.mako:22-33
template <int NS, int EOS>
__device__ __forceinline__ void frompT(const MixMom<NS, EOS> &m, double p, double T, EosState &s) {
#pragma clang fp contract(off)
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
    ideal_cph<NS>(m, T, cp, h);
    s.e = h - m.R * T;
    s.cp = cp;
    s.gamma = cp / (cp - m.R);
  }
  s.c = sqrt(s.gamma * m.R * T);
}

__device__ __forceinline__ void store_state(const View &st, int64_t o, const EosState &s) {
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
  fromrhoE<NS, EOS>(m, rho, e, st.at(o, 1), s);
  store_state(st, o, s);
}

}  // namespace pga
