// The alpha-damped viscous flux of one face (navstokes/kernels/flux/alpha-damping.mako:3-99), in
// registers: shared by the viscous sweep of one axis (pga_diffflux) and the launch over every face of
// a block (pga_faces), so both give the same bits.
#pragma once

#include "pga_common.hpp"

namespace pga {

// navstokes/kernels/flux/alpha-damping.mako:3-99: the alpha-damped viscous flux of face (i, j, k) of
// `axis` added to its flux Fv (every component but the mass), in registers
template <int NS>
__device__ __forceinline__ void diff_face(const View &Q, const View &st, const View &tr, const View &hsp,
                                          const View &gr, const View &xcc, const View &xf, const View &S, double alpha,
                                          int axis, int i, int j, int k, double (&Fv)[NS + 4]) {
#pragma clang fp contract(off)
  const int64_t oR = Q.off(i, j, k), oL = oR - (axis == 0 ? 1 : axis == 1 ? Q.sj : Q.sk);
  const int64_t ofa = S.off(i, j, k), oxf = xf.off(i, j, k);
  const double bulkvisc = 0.0;
  const double Sx = S.at(ofa, 0), Sy = S.at(ofa, 1), Sz = S.at(ofa, 2);
  const double Sabs = sqrt(Sx * Sx + Sy * Sy + Sz * Sz), Sinv = 1.0 / Sabs;
  const double nx = Sx * Sinv, ny = Sy * Sinv, nz = Sz * Sinv;
  const double mu = 0.5 * (tr.at(oR, 0) + tr.at(oL, 0));
  const double kappa = 0.5 * (tr.at(oR, 1) + tr.at(oL, 1));
  const double lambda = bulkvisc - 2.0 / 3.0 * mu;
  const double rhoinvL = st.at(oL, 2), rhoinvR = st.at(oR, 2);
  const double uL = Q.at(oL, 1) * rhoinvL, vL = Q.at(oL, 2) * rhoinvL, wL = Q.at(oL, 3) * rhoinvL;
  const double uR = Q.at(oR, 1) * rhoinvR, vR = Q.at(oR, 2) * rhoinvR, wR = Q.at(oR, 3) * rhoinvR;
  const double ex = xcc.at(oR, 0) - xcc.at(oL, 0), ey = xcc.at(oR, 1) - xcc.at(oL, 1), ez = xcc.at(oR, 2) - xcc.at(oL, 2);
  const double en0 = ex * nx, en1 = ey * ny, en2 = ez * nz;
  const double dampinv = alpha / sqrt(en0 * en0 + en1 * en1 + en2 * en2);
  const double dampx = dampinv * nx, dampy = dampinv * ny, dampz = dampinv * nz;
  const double xLx = xf.at(oxf, 0) - xcc.at(oL, 0), xLy = xf.at(oxf, 1) - xcc.at(oL, 1), xLz = xf.at(oxf, 2) - xcc.at(oL, 2);
  const double xRx = xf.at(oxf, 0) - xcc.at(oR, 0), xRy = xf.at(oxf, 1) - xcc.at(oR, 1), xRz = xf.at(oxf, 2) - xcc.at(oR, 2);
  const auto facegrad = [&](int l, double phiL, double phiR, double &gx, double &gy, double &gz) {
    const double gLx = gr.at(oL, 3 * l), gLy = gr.at(oL, 3 * l + 1), gLz = gr.at(oL, 3 * l + 2);
    const double gRx = gr.at(oR, 3 * l), gRy = gr.at(oR, 3 * l + 1), gRz = gr.at(oR, 3 * l + 2);
    const double wl = phiL + gLx * xLx + gLy * xLy + gLz * xLz;
    const double wr = phiR + gRx * xRx + gRy * xRy + gRz * xRz;
    const double jump = wr - wl;
    gx = 0.5 * (gRx + gLx) + dampx * jump;
    gy = 0.5 * (gRy + gLy) + dampy * jump;
    gz = 0.5 * (gRz + gLz) + dampz * jump;
  };
  double dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz;
  facegrad(0, uL, uR, dudx, dudy, dudz);
  facegrad(1, vL, vR, dvdx, dvdy, dvdz);
  facegrad(2, wL, wR, dwdx, dwdy, dwdz);
  const double div = dudx + dvdy + dwdz;
  const double txx = -2.0 * mu * dudx - lambda * div;
  const double txy = -mu * (dvdx + dudy);
  const double txz = -mu * (dwdx + dudz);
  const double tyy = -2.0 * mu * dvdy - lambda * div;
  const double tyz = -mu * (dwdy + dvdz);
  const double tzz = -2.0 * mu * dwdz - lambda * div;
  Fv[1] += txx * Sx + txy * Sy + txz * Sz;
  Fv[2] += txy * Sx + tyy * Sy + tyz * Sz;
  Fv[3] += txz * Sx + tyz * Sy + tzz * Sz;
  double dTx, dTy, dTz;
  facegrad(3, st.at(oL, 1), st.at(oR, 1), dTx, dTy, dTz);
  const double heatflux = -kappa * (dTx * Sx + dTy * Sy + dTz * Sz);
  const double uf = 0.5 * (uR + uL), vf = 0.5 * (vR + vL), wf = 0.5 * (wR + wL);
  double f4 = (uf * txx + vf * txy + wf * txz) * Sx + (uf * txy + vf * tyy + wf * tyz) * Sy +
              (uf * txz + vf * tyz + wf * tzz) * Sz + heatflux;
  double YL[NS], YR[NS];
#pragma unroll
  for (int n = 0; n < NS - 1; n++) {
    YL[n] = Q.at(oL, 5 + n) * rhoinvL;
    YR[n] = Q.at(oR, 5 + n) * rhoinvR;
  }
  double Dk, Vc = 0.0, gradYns = 0.0, sumYh = 0.0, Yns = 1.0;
  const double rho = 0.5 * (Q.at(oR, 0) + Q.at(oL, 0));
  double Fsp[NS];
#pragma unroll
  for (int n = 0; n < NS - 1; n++) {
    Dk = 0.5 * (tr.at(oR, 2 + n) + tr.at(oL, 2 + n));
    double dYx, dYy, dYz;
    facegrad(4 + n, YL[n], YR[n], dYx, dYy, dYz);
    const double gradYk = dYx * Sx + dYy * Sy + dYz * Sz;
    gradYns -= gradYk;
    Vc += Dk * gradYk;
    const double Jk = -rho * Dk * gradYk;
    const double hk = 0.5 * (hsp.at(oR, n) + hsp.at(oL, n));
    const double Yk = 0.5 * (YR[n] + YL[n]);
    Yns -= Yk;
    sumYh += Yk * hk;
    Fsp[n] = Fv[5 + n] + Jk;
    f4 += Jk * hk;
  }
  Dk = 0.5 * (tr.at(oR, 2 + NS - 1) + tr.at(oL, 2 + NS - 1));
  Vc += Dk * gradYns;
  const double corr = rho * Vc;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Fv[5 + n] = Fsp[n] + 0.5 * (YR[n] + YL[n]) * corr;
  f4 += corr * sumYh;
  const double hns = 0.5 * (hsp.at(oR, NS - 1) + hsp.at(oL, NS - 1));
  f4 += (-rho * Dk * gradYns + Yns * corr) * hns;
  Fv[4] += f4;
}

}  // namespace pga
