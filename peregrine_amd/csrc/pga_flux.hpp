// The advective face flux on the device: face sides by constant or MUSCL
// reconstruction, then KEEP (kepaec), HLLC or Rusanov.
//
// Every function here is evaluated with one rounding per written operation, in the order the
// reference's templates write it: the face fluxes feed the next stage's EOS, whose ln T
// polynomials turn any last-bit difference into 1e-11 of T, so a contracted multiply-add
// here would show at 1e-10 after a hundred steps.
#pragma once

#include "pga_eos.hpp"

namespace pga {

enum { RECON_CONSTANT = 0, RECON_MUSCL = 1 };
enum { FLUX_KEPAEC = 0, FLUX_HLLC = 1, FLUX_RUSANOV = 2 };

template <int NS>
struct FaceState {
  EosState s;
  double u, v, w;
  double Y[NS > 1 ? NS - 1 : 1];
};

enum { LIM_NONE = 0, LIM_KOREN = 1, LIM_MINMOD = 2, LIM_MC = 3, LIM_SUPERBEE = 4, LIM_VANLEER = 5 };

// b phi(a / b), b > 0, without the ratio (recon/{koren,minmod,mc,superbee,vanleer}.mako)
template <int LIM>
__device__ __forceinline__ double limit(double a, double b) {
#pragma clang fp contract(off)
  if constexpr (LIM == LIM_KOREN) return fmax(0.0, fmin(fmin(2.0 * a, (2.0 * b + a) * (1.0 / 3.0)), 2.0 * b));
  if constexpr (LIM == LIM_MINMOD) return fmax(0.0, fmin(b, a));
  if constexpr (LIM == LIM_MC) return fmax(0.0, fmin(fmin(2.0 * a, (b + a) * 0.5), 2.0 * b));
  if constexpr (LIM == LIM_SUPERBEE) return fmax(0.0, fmax(fmin(2.0 * a, b), fmin(a, 2.0 * b)));
  return (a + fabs(a)) * b / (b + fabs(a));
}

// recon/muscl.mako:5-13: the limited slope across the face carried to dp's sign and back; zero
// where dp is. Evaluated on every lane and selected, so the 26 slopes of a face stay one basic
// block the scheduler can interleave (a branch per slope would fence each one off)
template <int LIM>
__device__ __forceinline__ double phi(double dm, double dp) {
#pragma clang fp contract(off)
  const double s = copysign(1.0, dp);
  const double p = s * limit<LIM>(s * dm, fabs(dp));
  return dp != 0.0 ? p : 0.0;
}

template <int LIM>
__device__ __forceinline__ double left_of(double vLL, double vL, double vR) {
#pragma clang fp contract(off)
  return vL + 0.5 * phi<LIM>(vL - vLL, vR - vL);
}

template <int LIM>
__device__ __forceinline__ double right_of(double vL, double vR, double vRR) {
#pragma clang fp contract(off)
  return vR - 0.5 * phi<LIM>(vRR - vR, vR - vL);
}

// the face's sides: recon/constant.mako:3-18 (the two cells as they are) or
// recon/muscl.mako:15-51 (p, T, Y and velocity limited, each side closed by the eos from p, T, Y);
// o is the offset of the face's right cell R, d the stride along the axis
// evaluated with one rounding per operation: the sides' p and T feed the EOS's ill-conditioned
// ln T polynomials, where a last-bit change in T moves e by 1e-11 of itself
template <int NS, int EOS, int RECON, int LIM>
__device__ __forceinline__ void face_sides(const Gas<NS> &g, const View &Q, const View &st, int64_t o, int64_t d,
                                           FaceState<NS> &L, FaceState<NS> &R) {
#pragma clang fp contract(off)
  if constexpr (RECON == RECON_CONSTANT) {
    const int64_t oL = o - d;
    load_state(st, oL, L.s);
    load_state(st, o, R.s);
    const double rL = Q.at(oL, 0), rR = Q.at(o, 0);
    L.u = Q.at(oL, 1) / rL;
    L.v = Q.at(oL, 2) / rL;
    L.w = Q.at(oL, 3) / rL;
    R.u = Q.at(o, 1) / rR;
    R.v = Q.at(o, 2) / rR;
    R.w = Q.at(o, 3) / rR;
#pragma unroll
    for (int n = 0; n < NS - 1; n++) {
      L.Y[n] = Q.at(oL, 5 + n) / rL;
      R.Y[n] = Q.at(o, 5 + n) / rR;
    }
  } else {
    const int64_t oLL = o - 2 * d, oL = o - d, oRR = o + d;
    const double ri0 = st.at(oLL, 2), ri1 = st.at(oL, 2), ri2 = st.at(o, 2), ri3 = st.at(oRR, 2);
    const double p0 = st.at(oLL, 0), p1 = st.at(oL, 0), p2 = st.at(o, 0), p3 = st.at(oRR, 0);
    const double T0 = st.at(oLL, 1), T1 = st.at(oL, 1), T2 = st.at(o, 1), T3 = st.at(oRR, 1);
    const double pL = left_of<LIM>(p0, p1, p2), pR = right_of<LIM>(p1, p2, p3);
    const double TL = left_of<LIM>(T0, T1, T2), TR = right_of<LIM>(T1, T2, T3);
#pragma unroll
    for (int n = 0; n < NS - 1; n++) {
      const double y0 = Q.at(oLL, 5 + n) * ri0, y1 = Q.at(oL, 5 + n) * ri1;
      const double y2 = Q.at(o, 5 + n) * ri2, y3 = Q.at(oRR, 5 + n) * ri3;
      L.Y[n] = left_of<LIM>(y0, y1, y2);
      R.Y[n] = right_of<LIM>(y1, y2, y3);
    }
    MixMom<NS, EOS> mL, mR;
    mixmom2<NS, EOS>(g, [&](int n) { return L.Y[n]; }, [&](int n) { return R.Y[n]; }, mL, mR);
    frompT<NS, EOS>(g, mL, [&](int n) { return L.Y[n]; }, pL, TL, L.s);
    double vel[3][4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      vel[c][0] = Q.at(oLL, 1 + c) * ri0;
      vel[c][1] = Q.at(oL, 1 + c) * ri1;
      vel[c][2] = Q.at(o, 1 + c) * ri2;
      vel[c][3] = Q.at(oRR, 1 + c) * ri3;
    }
    L.u = left_of<LIM>(vel[0][0], vel[0][1], vel[0][2]);
    L.v = left_of<LIM>(vel[1][0], vel[1][1], vel[1][2]);
    L.w = left_of<LIM>(vel[2][0], vel[2][1], vel[2][2]);
    frompT<NS, EOS>(g, mR, [&](int n) { return R.Y[n]; }, pR, TR, R.s);
    R.u = right_of<LIM>(vel[0][1], vel[0][2], vel[0][3]);
    R.v = right_of<LIM>(vel[1][1], vel[1][2], vel[1][3]);
    R.w = right_of<LIM>(vel[2][1], vel[2][2], vel[2][3]);
  }
}

// function/kepaec.mako:2-20 (Shima et al. 2021): arithmetic means, the harmonic mean of cv T
template <int NS>
__device__ __forceinline__ void kepaec(const FaceState<NS> &L, const FaceState<NS> &R, const double S[3],
                                       double F[NS + 4]) {
#pragma clang fp contract(off)
  const double uf = 0.5 * (R.u + L.u), vf = 0.5 * (R.v + L.v), wf = 0.5 * (R.w + L.w);
  const double U = S[0] * uf + S[1] * vf + S[2] * wf;
  const double pf = 0.5 * (R.s.p + L.s.p);
  const double rho = 0.5 * (R.s.rho + L.s.rho);
  const double C = rho * U;
  F[0] = C;
  F[1] = C * uf + pf * S[0];
  F[2] = C * vf + pf * S[1];
  F[3] = C * wf + pf * S[2];
  const double Kj = C * 0.5 * (R.u * L.u + R.v * L.v + R.w * L.w);
  const double Pj = 0.5 * (L.s.p * (R.u * S[0] + R.v * S[1] + R.w * S[2]) + R.s.p * (L.u * S[0] + L.v * S[1] + L.w * S[2]));
  const double thL = L.s.cp / L.s.gamma * L.s.T, thR = R.s.cp / R.s.gamma * R.s.T;
  const double ef = 2.0 * thL * thR / (thL + thR) + 0.5 * ((L.s.e - thL) + (R.s.e - thR));
  F[4] = ef * C + Kj + Pj;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) F[5 + n] = 0.5 * (R.Y[n] + L.Y[n]) * C;
}

// function/hllc.mako:2-43: Davis wave speeds, the star jump through drho and dU
template <int NS>
__device__ __forceinline__ void hllc(const FaceState<NS> &L, const FaceState<NS> &R, const double S[3],
                                     double F[NS + 4]) {
#pragma clang fp contract(off)
  const double area = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  const double nx = S[0] / area, ny = S[1] / area, nz = S[2] / area;
  const double unL = (L.u * S[0] + L.v * S[1] + L.w * S[2]) / area;
  const double unR = (R.u * S[0] + R.v * S[1] + R.w * S[2]) / area;
  const double SL = unL - L.s.c, SR = unR + R.s.c;
  const double Sstar = (R.s.p - L.s.p + L.s.rho * unL * (SL - unL) - R.s.rho * unR * (SR - unR)) /
                       (L.s.rho * (SL - unL) - R.s.rho * (SR - unR));
  const bool left = SL >= 0.0 || Sstar >= 0.0;
  const bool star = !(SL >= 0.0) && (Sstar >= 0.0 || SR >= 0.0);
  // the upwind side's values, selected one by one (a reference to one side would put both in memory)
  const double Xrho = left ? L.s.rho : R.s.rho, Xp = left ? L.s.p : R.s.p, Xe = left ? L.s.e : R.s.e;
  const double Xu = left ? L.u : R.u, Xv = left ? L.v : R.v, Xw = left ? L.w : R.w;
  const double U = left ? unL : unR, SX = left ? SL : SR;
  const double et = Xe + 0.5 * (Xu * Xu + Xv * Xv + Xw * Xw);
  const double mdot = U * Xrho * area;
  F[0] = mdot;
  F[1] = mdot * Xu + Xp * S[0];
  F[2] = mdot * Xv + Xp * S[1];
  F[3] = mdot * Xw + Xp * S[2];
  F[4] = mdot * et + U * Xp * area;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) F[5 + n] = mdot * (left ? L.Y[n] : R.Y[n]);
  if (star) {
    const double rhostar = Xrho * ((SX - U) / (SX - Sstar));
    const double drho = rhostar - Xrho, dU = rhostar * (Sstar - U);
    const double jump = SX * area;
    F[0] += jump * drho;
    F[1] += jump * (drho * Xu + dU * nx);
    F[2] += jump * (drho * Xv + dU * ny);
    F[3] += jump * (drho * Xw + dU * nz);
    F[4] += jump * (drho * et + dU * (Sstar + Xp / (Xrho * (SX - U))));
#pragma unroll
    for (int n = 0; n < NS - 1; n++) F[5 + n] += jump * drho * (left ? L.Y[n] : R.Y[n]);
  }
}

// function/rusanov.mako:2-19: the central flux less the largest Davis wave speed's jump
template <int NS>
__device__ __forceinline__ void rusanov(const FaceState<NS> &L, const FaceState<NS> &R, const double S[3],
                                        double F[NS + 4]) {
#pragma clang fp contract(off)
  const double area = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  const double unL = (L.u * S[0] + L.v * S[1] + L.w * S[2]) / area;
  const double unR = (R.u * S[0] + R.v * S[1] + R.w * S[2]) / area;
  const double lam = fmax(fabs(unL) + L.s.c, fabs(unR) + R.s.c) * area;
  const double UL = unL * area, UR = unR * area;
  const double rL = L.s.rho, rR = R.s.rho;
  F[0] = 0.5 * (UR * rR + UL * rL - lam * (rR - rL));
  F[1] = 0.5 * (UR * rR * R.u + R.s.p * S[0] + UL * rL * L.u + L.s.p * S[0] - lam * (rR * R.u - rL * L.u));
  F[2] = 0.5 * (UR * rR * R.v + R.s.p * S[1] + UL * rL * L.v + L.s.p * S[1] - lam * (rR * R.v - rL * L.v));
  F[3] = 0.5 * (UR * rR * R.w + R.s.p * S[2] + UL * rL * L.w + L.s.p * S[2] - lam * (rR * R.w - rL * L.w));
  const double EL = rL * (L.s.e + 0.5 * (L.u * L.u + L.v * L.v + L.w * L.w));
  const double ER = rR * (R.s.e + 0.5 * (R.u * R.u + R.v * R.v + R.w * R.w));
  F[4] = 0.5 * (UR * (ER + R.s.p) + UL * (EL + L.s.p) - lam * (ER - EL));
#pragma unroll
  for (int n = 0; n < NS - 1; n++) {
    const double rhoYL = rL * L.Y[n], rhoYR = rR * R.Y[n];
    F[5 + n] = 0.5 * (rhoYR * UR + rhoYL * UL - lam * (rhoYR - rhoYL));
  }
}

template <int NS, int PRIMARY>
__device__ __forceinline__ void flux_function(const FaceState<NS> &L, const FaceState<NS> &R, const double S[3],
                                              double F[NS + 4]) {
  if constexpr (PRIMARY == FLUX_KEPAEC)
    kepaec<NS>(L, R, S, F);
  else if constexpr (PRIMARY == FLUX_HLLC)
    hllc<NS>(L, R, S, F);
  else
    rusanov<NS>(L, R, S, F);
}

// the face (i, j, k) of `axis` (between cells k - 1 and k along it): its sides, its flux into Fv
template <int NS, int EOS, int RECON, int PRIMARY, int LIM>
__device__ __forceinline__ void flux_at(const double *__restrict__ K, int axis, const View &Q, const View &st,
                                        const View &S, int i, int j, int k, double (&Fv)[NS + 4]) {
  const Gas<NS> g{K};
  const int64_t o = Q.off(i, j, k);
  const int64_t d = axis == 0 ? 1 : axis == 1 ? Q.sj : Q.sk;
  FaceState<NS> L, R;
  face_sides<NS, EOS, RECON, LIM>(g, Q, st, o, d, L, R);
  // the face's area vector is read by this face alone
  const int64_t of = S.off(i, j, k);
  const double Sv[3] = {__builtin_nontemporal_load(&S.at(of, 0)), __builtin_nontemporal_load(&S.at(of, 1)),
                        __builtin_nontemporal_load(&S.at(of, 2))};
  flux_function<NS, PRIMARY>(L, R, Sv, Fv);
}

// the face's flux into the face array
template <int NS, int EOS, int RECON, int PRIMARY, int LIM>
__device__ __forceinline__ void flux_face(const double *__restrict__ K, int axis, const View &Q, const View &st,
                                          const View &S, const View &F, int i, int j, int k) {
  double Fv[NS + 4];
  flux_at<NS, EOS, RECON, PRIMARY, LIM>(K, axis, Q, st, S, i, j, k, Fv);
  const int64_t oF = F.off(i, j, k);
#pragma unroll
  for (int l = 0; l < NS + 4; l++) __builtin_nontemporal_store(Fv[l], &F.at(oF, l));
}

}  // namespace pga
