// Shared device definitions: array views, the case constant layout, launch shapes,
// and the table log/exp the finite-rate kernels use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/peregrine_amd.h"
#include "generated/pga_tables.hpp"

namespace pga {

// cppoly rows hold 7 terms (ln T refits of degree <= 6), gpoly rows 8
constexpr int NT = 7;
constexpr int NG = 8;
constexpr int MAXNS = 16;

// ---------------------------------------------------------------------------
// the case constant table: one flat array of doubles, sections per species
// ---------------------------------------------------------------------------
struct Offsets {
  int MW, MWinv, Rsp, cppolyR, h298R, cp0, gpoly, h298, mu0, kappa0, lewis, size;
};

__host__ __device__ constexpr Offsets offsets(int ns) {
  Offsets o{};
  o.MW = 0;
  o.MWinv = o.MW + ns;
  o.Rsp = o.MWinv + ns;
  o.cppolyR = o.Rsp + ns;
  o.h298R = o.cppolyR + ns * NT;
  o.cp0 = o.h298R + ns;
  o.gpoly = o.cp0 + ns;
  o.h298 = o.gpoly + ns * NG;
  o.mu0 = o.h298 + ns;
  o.kappa0 = o.mu0 + ns;
  o.lewis = o.kappa0 + ns;
  o.size = o.lewis + ns;
  return o;
}

// the species constants of a case with NS species, read at compile-time offsets from a
// wave-uniform pointer: the compiler turns each read into a scalar load
template <int NS>
struct Gas {
  static constexpr Offsets O = offsets(NS);
  const double *__restrict__ K;
  __device__ __forceinline__ double MW(int n) const { return K[O.MW + n]; }
  __device__ __forceinline__ double MWinv(int n) const { return K[O.MWinv + n]; }
  __device__ __forceinline__ double Rsp(int n) const { return K[O.Rsp + n]; }
  __device__ __forceinline__ double cppolyR(int n, int m) const { return K[O.cppolyR + n * NT + m]; }
  __device__ __forceinline__ double h298R(int n) const { return K[O.h298R + n]; }
  __device__ __forceinline__ double cp0(int n) const { return K[O.cp0 + n]; }
  __device__ __forceinline__ double gpoly(int n, int m) const { return K[O.gpoly + n * NG + m]; }
  __device__ __forceinline__ double h298(int n) const { return K[O.h298 + n]; }
};

// ---------------------------------------------------------------------------
// views and launch shapes
// ---------------------------------------------------------------------------
struct View {
  double *__restrict__ d;
  int64_t sj, sk, sc;
  __device__ __forceinline__ int64_t off(int i, int j, int k) const { return i + j * sj + k * sk; }
  __device__ __forceinline__ double &at(int64_t o, int c) const { return d[o + c * sc]; }
};

__host__ inline View view(const pga_view &v) { return View{v.data, v.sj, v.sk, v.sc}; }

// cells along i map to the 64 lanes of a wave, 4 waves step along j; blocks tile j and k
constexpr int TI = 64;
constexpr int TJ = 4;

struct Box3 {
  int s0, s1, s2, e0, e1, e2;
};

__host__ inline Box3 box3(const pga_box &b) {
  return Box3{b.start[0], b.start[1], b.start[2], b.extent[0], b.extent[1], b.extent[2]};
}

__host__ inline dim3 grid_of(const Box3 &b) {
  return dim3((b.e0 + TI - 1) / TI, (b.e1 + TJ - 1) / TJ, b.e2);
}

// the cell a thread takes, false when the block overhangs the box
__device__ __forceinline__ bool cell_of(const Box3 &b, int &i, int &j, int &k) {
  const int ii = blockIdx.x * TI + threadIdx.x, jj = blockIdx.y * TJ + threadIdx.y;
  i = b.s0 + ii;
  j = b.s1 + jj;
  k = b.s2 + blockIdx.z;
  return ii < b.e0 && jj < b.e1;
}

// ---------------------------------------------------------------------------
// tlog / texp: log by a 128-bin table over the mantissa and a 5-term series,
// exp by the 64 powers 2^(j/64) and a 6-term series (the reference's table
// methods, backend/base/mathfns.py:1-22 and kernels/mathfns.mako:6-54)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tlog(double x) {
  long long bits = __double_as_longlong(x);
  const int expo = (int)((bits >> 52) & 0x7ff) - 1022;
  const int bin = (int)((bits >> (52 - LOGBITS)) & (LOGBINS - 1));
  bits = (bits & 0x000fffffffffffffLL) | 0x3fe0000000000000LL;
  const double mant = __longlong_as_double(bits);
  const double u = mant * RCPTAB[bin] - 1.0;
  double s = 0.2;
  s = s * u - 0.25;
  s = s * u + 1.0 / 3.0;
  s = s * u - 0.5;
  s = s * u + 1.0;
  return expo * TAB_LN2 + LOGTAB[bin] + u * s;
}

__device__ __forceinline__ double texp(double x) {
  const double t = x * TAB_EXPK + 6755399441055744.0;
  long long bits = __double_as_longlong(t);
  const int kexp = (int)bits, nexp = kexp >> EXPBITS;
  const double kd = t - 6755399441055744.0;
  const double r = (x - kd * TAB_LN2BINS_HI) - kd * TAB_LN2BINS_LO;
  double s = 1.0 / 120.0;
  s = s * r + 1.0 / 24.0;
  s = s * r + 1.0 / 6.0;
  s = s * r + 0.5;
  s = s * r + 1.0;
  s = s * r + 1.0;
  const double m = EXPTAB[kexp & (EXPBINS - 1)] * s;
  bits = __double_as_longlong(m);
  bits += (long long)(nexp > -1022 ? nexp : -1022) << 52;
  return __longlong_as_double(bits);
}

}  // namespace pga
