// Finite-rate chemistry on the device: the mass fractions a cell's rates are taken at
// and the coupled source that begins dQ. The production fold itself is generated per
// mechanism (generated/mech_*.hpp).
#pragma once

#include "pga_common.hpp"

namespace pga {

// g/(Ru T) of species n: its Gibbs polynomial in ln T plus h298/T (eos/kernels/species.mako:16-21),
// Horner steps as explicit fused multiply-adds (the oracle evaluates the same sequence): the
// polynomial's terms cancel by six orders, so the rounding sequence is fixed, not left to the compiler
template <int NS>
__device__ __forceinline__ double gibbs_rt(const Gas<NS> &g, int n, double logT, double Tinv) {
#pragma clang fp contract(off)
  double gr = 0.0;
#pragma unroll
  for (int m = NG - 1; m >= 0; m--) gr = fma(gr, logT, g.gpoly(n, m));
  return fma(g.h298(n), Tinv, gr);
}

// cp/R and the internal energy of species n at T (eos/kernels/species.mako:3-27), u = ln T
template <int NS>
__device__ __forceinline__ double cpR_species(const Gas<NS> &g, int n, double u) {
#pragma clang fp contract(off)
  double cpR = 0.0;
#pragma unroll
  for (int m = NT - 1; m >= 0; m--) cpR = cpR * u + g.cppoly(n, m);
  return cpR;
}

template <int NS>
__device__ __forceinline__ double e_species(const Gas<NS> &g, int n, double T, double u, double Tinv) {
#pragma clang fp contract(off)
  double hRT = 0.0;
#pragma unroll
  for (int m = NT - 1; m >= 0; m--) hRT = hRT * u + g.hpoly(n, m);
  hRT += g.h298(n) * Tinv;
  return (hRT - 1.0) * T * g.Rsp(n);
}

// euler/kernels/solution.mako:25-35: Y clamped to [0, 1], rescaled when their sum
// passes one, the last species what they leave (never below zero)
template <int NS>
__device__ __forceinline__ void massfrac(const View &Q, int64_t o, double rhoinv, double (&Y)[NS]) {
  double Ylast = 1.0, Ysum = 0.0;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) {
    const double Yn = fmax(fmin(Q.at(o, 5 + n) * rhoinv, 1.0), 0.0);
    Y[n] = Yn;
    Ylast -= Yn;
    Ysum += Yn;
  }
  const double Yscale = 1.0 / fmax(Ysum, 1.0);
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Y[n] *= Yscale;
  Y[NS - 1] = fmax(Ylast, 0.0);
}

// chemint/kernels/source.mako:3-10 and cell.mako:5-20 for the cell at o
template <class M>
__device__ __forceinline__ void cell_source(const Gas<M::NS> &g, const double *__restrict__ C, const View &Q,
                                            const View &st, const View &dQ, int64_t o) {
  constexpr int NS = M::NS;
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  double Y[NS], omega[NS];
  massfrac<NS>(Q, o, rhoinv, Y);
  M::production(g, C, rho, Y, st.at(o, 1), omega);
#pragma unroll
  for (int l = 0; l < 5; l++) dQ.at(o, l) = 0.0;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) dQ.at(o, 5 + n) = omega[n];
}

}  // namespace pga
