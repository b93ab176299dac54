// Shared device definitions: array views, the case constant layout, launch shapes,
// and the table log/exp the finite-rate kernels use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/peregrine_amd.h"
#include "generated/pga_tables.hpp"
#define PGA_ELOG_TABLE static __constant__ const
#include "generated/pga_elog_tables.h"

namespace pga {

// cppoly rows hold 7 terms (ln T refits of degree <= 6), gpoly rows 8
constexpr int NT = 7;
constexpr int NG = 8;
constexpr int MAXNS = 16;
// J/kmol/K, the universal gas constant as the mixture tables state it
constexpr double Ru = 8314.46261815324;

// ---------------------------------------------------------------------------
// the case constant table: one flat array of doubles, sections per species
// ---------------------------------------------------------------------------
struct Offsets {
  int MW, MWinv, Rsp, cppolyR, h298R, cp0, gpoly, h298, mu0, kappa0, lewis, hpoly, MWqinv, wilkepair, mupoly,
      kappapoly, cppoly, cubic, cubicsqrtac, cubicb, cubicda, cubicmTc, cubiconepm, cubicsqrtacMW, cubicbMW, vshift,
      vshiftMW, cubicdaMW, acentric, chungsigma, chungepsk, chungvisc, chungcond, pairfit, selffit, bathof, diffmeta,
      size;
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
  o.hpoly = o.lewis + ns;
  o.MWqinv = o.hpoly + ns * NT;
  o.wilkepair = o.MWqinv + ns;
  o.mupoly = o.wilkepair + ns * ns;
  o.kappapoly = o.mupoly + ns * NT;
  o.cppoly = o.kappapoly + ns * NT;
  // the cubic real gas: u, w, sigma, eps, then per species
  o.cubic = o.cppoly + ns * NT;
  o.cubicsqrtac = o.cubic + 4;
  o.cubicb = o.cubicsqrtac + ns;
  o.cubicda = o.cubicb + ns;
  o.cubicmTc = o.cubicda + ns;
  o.cubiconepm = o.cubicmTc + ns;
  o.cubicsqrtacMW = o.cubiconepm + ns;
  o.cubicbMW = o.cubicsqrtacMW + ns;
  o.vshift = o.cubicbMW + ns;
  o.vshiftMW = o.vshift + ns;
  o.cubicdaMW = o.vshiftMW + ns;
  o.acentric = o.cubicdaMW + ns;
  // Chung: sigma, epsilon/k per species, the E (10 x 4) and B (7 x 4) tables
  o.chungsigma = o.acentric + ns;
  o.chungepsk = o.chungsigma + ns;
  o.chungvisc = o.chungepsk + ns;
  o.chungcond = o.chungvisc + 40;
  // binary diffusion: each species' partner fits (ns x ns rows of NT, the first `width` used), its
  // own pair's fit, the partners' indices, and (width, partial bath)
  o.pairfit = o.chungcond + 28;
  o.selffit = o.pairfit + ns * ns * NT;
  o.bathof = o.selffit + ns * NT;
  o.diffmeta = o.bathof + ns * ns;
  o.size = o.diffmeta + 2;
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
  // transport: constant mu and kappa, Lewis numbers, Wilke's MW^-1/4 and pair factors, the
  // kinetic-theory fits of sqrt(mu) / T^1/4 and kappa / sqrt(T) in ln T; the enthalpy fits
  __device__ __forceinline__ double mu0(int n) const { return K[O.mu0 + n]; }
  __device__ __forceinline__ double kappa0(int n) const { return K[O.kappa0 + n]; }
  __device__ __forceinline__ double lewis(int n) const { return K[O.lewis + n]; }
  __device__ __forceinline__ double hpoly(int n, int m) const { return K[O.hpoly + n * NT + m]; }
  __device__ __forceinline__ double MWqinv(int n) const { return K[O.MWqinv + n]; }
  __device__ __forceinline__ double wilkepair(int n, int m) const { return K[O.wilkepair + n * NS + m]; }
  __device__ __forceinline__ double mupoly(int n, int m) const { return K[O.mupoly + n * NT + m]; }
  __device__ __forceinline__ double kappapoly(int n, int m) const { return K[O.kappapoly + n * NT + m]; }
  __device__ __forceinline__ double cppoly(int n, int m) const { return K[O.cppoly + n * NT + m]; }
  // the cubic: cubic(0..3) = u, w, sigma, eps
  __device__ __forceinline__ double cubic(int i) const { return K[O.cubic + i]; }
  __device__ __forceinline__ double cubicsqrtac(int n) const { return K[O.cubicsqrtac + n]; }
  __device__ __forceinline__ double cubicb(int n) const { return K[O.cubicb + n]; }
  __device__ __forceinline__ double cubicda(int n) const { return K[O.cubicda + n]; }
  __device__ __forceinline__ double cubicmTc(int n) const { return K[O.cubicmTc + n]; }
  __device__ __forceinline__ double cubiconepm(int n) const { return K[O.cubiconepm + n]; }
  __device__ __forceinline__ double cubicsqrtacMW(int n) const { return K[O.cubicsqrtacMW + n]; }
  __device__ __forceinline__ double cubicbMW(int n) const { return K[O.cubicbMW + n]; }
  __device__ __forceinline__ double vshift(int n) const { return K[O.vshift + n]; }
  __device__ __forceinline__ double vshiftMW(int n) const { return K[O.vshiftMW + n]; }
  __device__ __forceinline__ double cubicdaMW(int n) const { return K[O.cubicdaMW + n]; }
  __device__ __forceinline__ double acentric(int n) const { return K[O.acentric + n]; }
  __device__ __forceinline__ double chungsigma(int n) const { return K[O.chungsigma + n]; }
  __device__ __forceinline__ double chungepsk(int n) const { return K[O.chungepsk + n]; }
  __device__ __forceinline__ double chungvisc(int i, int j) const { return K[O.chungvisc + 4 * i + j]; }
  __device__ __forceinline__ double chungcond(int i, int j) const { return K[O.chungcond + 4 * i + j]; }
  __device__ __forceinline__ double pairfit(int n, int k, int t) const { return K[O.pairfit + (n * NS + k) * NT + t]; }
  __device__ __forceinline__ double selffit(int n, int t) const { return K[O.selffit + n * NT + t]; }
  __device__ __forceinline__ double bathof(int n, int k) const { return K[O.bathof + n * NS + k]; }
  __device__ __forceinline__ double diffmeta(int i) const { return K[O.diffmeta + i]; }
};

// ---------------------------------------------------------------------------
// views and launch shapes
// ---------------------------------------------------------------------------
struct View {
  double *__restrict__ d;
  int64_t sj, sk, sc;
  __device__ __forceinline__ int64_t off(int i, int j, int k) const { return i + j * sj + k * sk; }
  // the component's plane from the uniform base, then a 32-bit byte offset into it: a load takes
  // the scalar base and one vector offset, with no 64-bit address arithmetic per lane (every entry
  // point checks that a plane spans less than 4 GiB)
  __device__ __forceinline__ double &at(int64_t o, int c) const {
    return *reinterpret_cast<double *>(reinterpret_cast<char *>(d + c * sc) + (uint32_t)((uint64_t)o << 3));
  }
};

__host__ inline View view(const pga_view &v) { return View{v.data, v.sj, v.sk, v.sc}; }

// (x, y, z) turned by the row-major rotation R, each row's sum in the written order
__device__ __forceinline__ void turned(const double *R, double x, double y, double z, double &ox, double &oy,
                                       double &oz) {
#pragma clang fp contract(off)
  ox = R[0] * x + R[1] * y + R[2] * z;
  oy = R[3] * x + R[4] * y + R[5] * z;
  oz = R[6] * x + R[7] * y + R[8] * z;
}

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

// Flat launches: 256-thread blocks over the box's cells in i-fastest order, so a row
// of 129 faces costs 129 lanes rather than three 64-lane waves.  The hardware deals
// consecutive blocks round-robin to the 8 XCDs, each with its own L2; the block
// number is remapped so that every XCD walks one contiguous run of the box and the
// rows a stencil shares with its neighbours stay in one L2.
constexpr int FLAT = 256;
constexpr int NXCD = 8;

// a flat launch indexes its cells in 32 bits: boxes beyond FLAT_MAX cells go in k-slabs
constexpr int64_t FLAT_MAX = int64_t(1) << 31;

__host__ inline dim3 flat_grid(const Box3 &b) {
  return dim3((unsigned)(((int64_t)b.e0 * b.e1 * b.e2 + FLAT - 1) / FLAT));
}

// launch(sub-box) over slabs of whole k-planes, each under FLAT_MAX cells
template <class F>
__host__ inline void flat_slabs(const Box3 &b, F &&launch) {
  const int64_t plane = (int64_t)b.e0 * b.e1;
  const int per = (int)std::max<int64_t>(1, (FLAT_MAX - FLAT) / std::max<int64_t>(plane, 1));
  for (int k0 = 0; k0 < b.e2; k0 += per) {
    Box3 s = b;
    s.s2 = b.s2 + k0;
    s.e2 = std::min(per, b.e2 - k0);
    launch(s);
  }
}

__device__ __forceinline__ unsigned xcd_block() {
  const unsigned nb = gridDim.x, bid = blockIdx.x, x = bid % NXCD, q = nb / NXCD, r = nb % NXCD;
  return x * q + (x < r ? x : r) + bid / NXCD;
}

__device__ __forceinline__ bool flat_cell_of(const Box3 &b, int &i, int &j, int &k) {
  const unsigned t = xcd_block() * FLAT + threadIdx.x;
  const unsigned row = t / (unsigned)b.e0;
  i = b.s0 + (int)(t - row * (unsigned)b.e0);
  j = b.s1 + (int)(row % (unsigned)b.e1);
  k = b.s2 + (int)(row / (unsigned)b.e1);
  return row < (unsigned)b.e1 * (unsigned)b.e2;
}

// p (wave-uniform), as a value the compiler cannot see through: the loads behind it stay in the
// loop that makes them rather than being hoisted out of it (a mechanism's or a gas's hundreds of
// constants would otherwise be held in registers across the loop, at one wave per SIMD)
template <class T>
__device__ __forceinline__ T *opaque(T *p) {
  asm volatile("" : "+s"(p));
  return p;
}

// the cell of a flat tile t of box b, false past the box
__device__ __forceinline__ bool flat_at(const Box3 &b, int64_t t, int &i, int &j, int &k) {
  const int64_t c = t * FLAT + threadIdx.x;
  const int64_t row = c / b.e0;
  i = b.s0 + (int)(c - row * b.e0);
  j = b.s1 + (int)(row % b.e1);
  k = b.s2 + (int)(row / b.e1);
  return row < (int64_t)b.e1 * b.e2;
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
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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

// ---------------------------------------------------------------------------
// elog: the EOS's ln T. ln x = e ln 2 - ln r_k + ln(1 + (m r_k - 1)) for x = m 2^e, m in [1, 2),
// r_k a 9-bit reciprocal of the centre of m's bin by its top 7 bits; m r_k - 1 by one fused
// multiply-add, ln(1 + r) by its series to r^7, -ln r_k and ln 2 in two parts. IEEE operations
// and explicit fma only, so every platform gives the same bits (within an ulp of ln x): the ln T
// polynomials of the tpg EOS turn a last-bit change of ln T into 1e-11 of h.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double elog(double x) {
#pragma clang fp contract(off)
  // Every lane evaluates the table path, the series near 1 and the limits, and selects: the EOS
  // calls this on both sides of a face, and without branches the two calls and the code around them
  // stay one block the scheduler can interleave. A subnormal is scaled into the normal range;
  // zero, negative, NaN and infinity take their limits. No library call: a call's save and restore
  // around the EOS costs registers on every path
  const bool sub = x < 2.2250738585072014e-308;
  const double xs = sub ? x * 18014398509481984.0 : x;
  const long long bits = __double_as_longlong(xs);
  const int e = (int)((bits >> 52) & 0x7ff) - 1023 - (sub ? 54 : 0);
  const int k = (int)((bits >> 45) & 127);
  const double m = __longlong_as_double((bits & 0x000fffffffffffffLL) | 0x3ff0000000000000LL);
  const double r = fma(m, ELOG_RCP[k], -1.0);
  double q = 1.0 / 7.0;
  q = fma(q, r, -1.0 / 6.0);
  q = fma(q, r, 0.2);
  q = fma(q, r, -0.25);
  q = fma(q, r, 1.0 / 3.0);
  q = fma(q, r, -0.5);
  const double p = (r * r) * q;
  const double hi = e * PGA_ELOG_LN2HI + ELOG_LHI[k];
  const double lo = e * PGA_ELOG_LN2LO + ELOG_LLO[k];
  const double table = hi + (r + (p + lo));
  // near 1 the table's terms would cancel: ln(1 + r) by its series, r = x - 1 exact
  const double r1 = x - 1.0;
  double q1 = -1.0 / 14.0;
#pragma unroll
  for (int n = 13; n >= 2; n--) q1 = fma(q1, r1, ((n & 1) ? 1.0 : -1.0) / n);
  const double series = r1 + (r1 * r1) * q1;
  const double finite = fabs(r1) < 0.03125 ? series : table;
  const double limit = x == 0.0 ? -__builtin_huge_val() : x > 0.0 ? x : __builtin_nan("");
  return x > 0.0 && x <= 1.7976931348623157e308 ? finite : limit;
}

// ln(1 + x) through elog: u = 1 + x rounded, then the first-order correction for that rounding,
// ((u - 1) - x) / u; one specified method, so the device and the oracle give the same bits where the
// libraries' log1p differ in the last
__device__ __forceinline__ double dlog1p(double x) {
#pragma clang fp contract(off)
  const double u = 1.0 + x;
  return elog(u) - ((u - 1.0) - x) / u;
}

// ---------------------------------------------------------------------------
// dcbrt: the cube root by an exponent split and Newton steps in plain IEEE arithmetic, so the oracle
// and the device give the same bits (the libraries' cbrt round apart by an ulp, and the cubic EOS's
// root amplifies that across a pseudo-critical layer). x = a 2^(3q) with a in [0.5, 4); six Newton
// steps from a linear start settle within an ulp of cbrt(a); the 2^q is exact.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dcbrt(double x) {
#pragma clang fp contract(off)
  if (x == 0.0 || !(fabs(x) <= 1.7976931348623157e308)) return x;
  int e;
  const double m = frexp(fabs(x), &e);
  int q = e / 3, r = e - 3 * q;
  if (r < 0) r += 3, q -= 1;
  const double a = ldexp(m, r);
  double y = 0.75 + 0.25 * a;
  for (int k = 0; k < 6; k++) y = y - (y * y * y - a) / (3.0 * (y * y));
  return copysign(ldexp(y, q), x);
}

}  // namespace pga
