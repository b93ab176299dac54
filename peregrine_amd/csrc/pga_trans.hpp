// Transport properties of a mixture: the species' viscosity and conductivity (constant values,
// or the kinetic-theory fits in ln T), Wilke's mixing rule and the conductivity mean, and Chung's
// dense-gas factors. Arithmetic follows the oracle operation for operation (contraction off).
#pragma once

#include "pga_eos.hpp"

namespace pga {

enum { TRANS_CONSTANT = 0, TRANS_KINETIC = 1, TRANS_CHUNG = 2 };

// transport/kernels/chung.mako:6-67: Chung et al.'s dense-gas factors on the dilute mu and kappa,
// one-fluid mixing, without the dipole and association terms; Chung's units inside
template <int NS>
__device__ __forceinline__ void chung(const Gas<NS> &g, const double (&X)[NS], double T, double lnT, double rho,
                                      double &mu, double &kappa) {
#pragma clang fp contract(off)
  double sigma3 = 0.0, epssigma3 = 0.0, masssum = 0.0, omegasigma3 = 0.0, cvR = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    double cpR = 0.0;
#pragma unroll
    for (int m = NT - 1; m >= 0; m--) cpR = cpR * lnT + g.cppoly(i, m);
    cvR += X[i] * (cpR - 1.0);
#pragma unroll
    for (int j = 0; j < NS; j++) {
      const double xx = X[i] * X[j];
      const double sigma2ij = g.chungsigma(i) * g.chungsigma(j);
      const double sigma3ij = sigma2ij * sqrt(sigma2ij);
      const double epsij = sqrt(g.chungepsk(i) * g.chungepsk(j));
      const double Mij = 2.0 * g.MW(i) * g.MW(j) / (g.MW(i) + g.MW(j));
      sigma3 += xx * sigma3ij;
      epssigma3 += xx * epsij * sigma3ij;
      masssum += xx * epsij * sigma2ij * sqrt(Mij);
      omegasigma3 += xx * 0.5 * (g.acentric(i) + g.acentric(j)) * sigma3ij;
    }
  }
  const double epsk = epssigma3 / sigma3;
  const double cbrtsigma3 = dcbrt(sigma3);
  const double sigma2 = cbrtsigma3 * cbrtsigma3;
  const double sqrtMm = masssum / (epsk * sigma2);
  const double Mm = sqrtMm * sqrtMm;
  const double omega = omegasigma3 / sigma3;
  const double Tc = 1.2593 * epsk;
  const double Vc = sigma3 / (0.809 * 0.809 * 0.809);
  const double Fc = 1.0 - 0.2756 * omega;
  const double cbrtVc = dcbrt(Vc);
  const double Vc23 = cbrtVc * cbrtVc;
  const double Mprime = Mm * 1.0e-3;
  double E[10], B[7];
#pragma unroll
  for (int i = 0; i < 10; i++) E[i] = g.chungvisc(i, 0) + g.chungvisc(i, 1) * omega;
#pragma unroll
  for (int i = 0; i < 7; i++) B[i] = g.chungcond(i, 0) + g.chungcond(i, 1) * omega;
  const double Tr = T / Tc;
  const double Tstar = T / epsk, invTstar = epsk / T;
  // pow and exp by the table methods, which the oracle evaluates to the same bits
  const double Omegav =
      1.16145 * texp(-0.14874 * elog(Tstar)) + 0.52487 * texp(-0.77320 * Tstar) + 2.16178 * texp(-2.43787 * Tstar);
  const double Yy = rho / Mm * 1.0e-3 * Vc / 6.0;
  const double omYy = 1.0 - Yy;
  const double G1 = (1.0 - 0.5 * Yy) / (omYy * omYy * omYy);
  const double G2 = (E[0] * (1.0 - texp(-E[3] * Yy)) / Yy + E[1] * G1 * texp(E[4] * Yy) + E[2] * G1) / (E[0] * E[3] + E[1] + E[2]);
  const double etastarstar = E[6] * Yy * Yy * G2 * texp(E[7] + E[8] * invTstar + E[9] * invTstar * invTstar);
  const double muratio = (1.0 / G2 + E[5] * Yy) + etastarstar * Omegav / (sqrt(Tstar) * Fc);
  mu = mu * muratio;
  const double eta0 = 4.0785e-6 * sqrt(Mm * T) / (Vc23 * Omegav) * Fc;
  const double alpha = cvR - 1.5;
  const double beta = 0.7862 - 0.7109 * omega + 1.3168 * omega * omega;
  const double Zc = 2.0 + 10.5 * Tr * Tr;
  const double Psi = 1.0 + alpha * ((0.215 + 0.28288 * alpha - 1.061 * beta + 0.26665 * Zc) / (0.6366 + beta * Zc + 1.061 * alpha * beta));
  const double lambda0 = 31.2 * eta0 * Psi / Mprime;
  const double H2 = (B[0] * (1.0 - texp(-B[3] * Yy)) / Yy + B[1] * G1 * texp(B[4] * Yy) + B[2] * G1) / (B[0] * B[3] + B[1] + B[2]);
  const double kapparatio = (1.0 / H2 + B[5] * Yy) + 3.586e-3 * sqrt(Tc / Mprime) / Vc23 * B[6] * Yy * Yy * H2 * sqrt(Tr) / lambda0;
  kappa = kappa * kapparatio;
}

// transport/kernels/{constant,dilute}.mako, mixrule/kernels/wilke.mako:3-17,
// transport/kernels/mixing.mako:3-10: mu and kappa of the mixture at T (lnT its log) and rho
template <int NS, int TRANS>
__device__ __forceinline__ void mu_kappa(const Gas<NS> &g, const double (&X)[NS], double T, double lnT, double rho,
                                         double &mu, double &kappa) {
#pragma clang fp contract(off)
  double sqrtmu[NS], kappasp[NS];
  if constexpr (TRANS == TRANS_CONSTANT) {
#pragma unroll
    for (int n = 0; n < NS; n++) sqrtmu[n] = sqrt(g.mu0(n)), kappasp[n] = g.kappa0(n);
  } else {
    const double sqrtT = sqrt(T), sqrtsqrtT = sqrt(sqrtT);
#pragma unroll
    for (int n = 0; n < NS; n++) {
      double fit = 0.0;
#pragma unroll
      for (int m = NT - 1; m >= 0; m--) fit = fit * lnT + g.mupoly(n, m);
      sqrtmu[n] = fit * sqrtsqrtT;
    }
#pragma unroll
    for (int n = 0; n < NS; n++) {
      double fit = 0.0;
#pragma unroll
      for (int m = NT - 1; m >= 0; m--) fit = fit * lnT + g.kappapoly(n, m);
      kappasp[n] = fit * sqrtT;
    }
  }
  double sinv[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) sinv[n] = 1.0 / (sqrtmu[n] * g.MWqinv(n));
  mu = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const double sn = sqrtmu[n] * g.MWqinv(n);
    double phi = 0.0;
#pragma unroll
    for (int m = 0; m < NS; m++) {
      const double r = 1.0 + sn * sinv[m];
      phi += r * r * g.wilkepair(n, m) * X[m];
    }
    mu += sqrtmu[n] * sqrtmu[n] * X[n] / phi;
  }
  double series = 0.0, parallel = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    series += X[n] * kappasp[n];
    parallel += X[n] / kappasp[n];
  }
  kappa = 0.5 * (series + 1.0 / parallel);
  if constexpr (TRANS == TRANS_CHUNG) chung(g, X, T, lnT, rho, mu, kappa);
}

}  // namespace pga
