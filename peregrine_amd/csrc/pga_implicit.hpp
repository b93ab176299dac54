// Implicit chemistry (device code, shared by pga_implicit.hip and the launches over a table of
// blocks in pga_batch.hip): each cell integrates its own finite-rate chemistry over the step by the
// linearly implicit (Rosenbrock) Euler method in `parts` equal parts, the same 0D system the
// sub-stepped form integrates explicitly (chemint/kernels/substep.mako:5-83: the ns mass fractions
// and T at constant density and internal energy, dT/dt = -sum e_n w_n / (rho cv)), and begins dQ
// with the mean rate over the step.
//
// A part of length h from y0 = (Y_0 .. Y_{ns-1}, T) takes y1 = y0 + h x with (I - h J) x = f(y0):
// backward Euler's first Newton step, first order and stable however stiff the chemistry. J is
// formed in FP64 by forward differences of f (one more evaluation of the rates per unknown); the
// matrix is kept in FP64 and, rounded to FP32, factored by a batched LU on the matrix cores. The
// solve is an iterative refinement: x += (LU)^-1 r with the residual r = f - A x taken in FP64
// against the FP64 matrix, until every row's residual is within `tolerance` of that row's scale
// (|f_i| + sum_j |A_ij x_j|, the componentwise backward error), so x is the FP64 solution whatever
// the factors' rounding. The Arrhenius exponentials, the Gibbs polynomials, the residuals and the
// update stay FP64.
//
// The batched LU takes 16 matrices per wave, padded to 12 x 12 (three 4 x 4 blocks a side), one
// lane quad per matrix: lane 4b + q holds column q of each of matrix b's blocks, a block's row i in
// its register i. That is the accumulator layout of v_mfma_f32_4x4x1f32 (16 independent 4 x 4
// products per instruction), so the Schur-complement updates of the blocked right-looking LU,
// A_IJ -= L_IK U_KJ, are four rank-one MFMAs each for all 16 matrices at once; the diagonal blocks'
// factorisation and the two triangular block solves run on the VALU within the quad, with DPP
// broadcasts. No pivoting: the matrices are I - h J of a chemical source, whose diagonal grows with
// the destruction rates; the FP64 refinement corrects what the FP32 factors leave.
//
// The implicit form's work array holds the FP64 matrices, [group of 16 cells][(ns + 1)^2 entries]
// [16 cells] (a row of the 16 cells' same entry is 128 contiguous bytes, so both the lane-per-cell
// kernels and the quad-per-cell solve move whole segments), then y, f at y's FP32 rounding and f at
// y, [ns + 1][cells] doubles each. The FP32 factors never leave the registers of the solve. The
// batched solve of caller-given systems keeps its FP32 factors in a work array the same way,
// [group][144 entries][16 cells].
#pragma once

#include <type_traits>

#include "generated/mechanisms.hpp"
#include "pga_chem.hpp"

namespace pga {

constexpr int IM_G = 16;                   // matrices per MFMA group: a wave's 16 lane quads
constexpr int IM_N = 12;                   // matrices padded to 12 x 12
constexpr int IM_E = IM_N * IM_N;          // entries of a padded matrix
constexpr int IM_GROUP = IM_E * IM_G;      // floats per group
constexpr int IM_FAIL = 1000;              // added to a cell's count for each part that did not converge

typedef float v4f __attribute__((ext_vector_type(4)));

// entry (r, c) of matrix `cell` in the group-major factor array
__device__ __forceinline__ int64_t lu_at(int64_t cell, int r, int c) {
  return (cell >> 4) * IM_GROUP + (int64_t)(r * IM_N + c) * IM_G + (cell & 15);
}

template <int B, int E, class F>
__host__ __device__ __forceinline__ void sfor(F &&f) {
  if constexpr (B < E) {
    f.template operator()<B>();
    sfor<B + 1, E>(f);
  }
}

// lane P's x within each quad, to all four lanes of the quad (DPP quad_perm P, P, P, P)
template <int P>
__device__ __forceinline__ float qb(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), P * 0x55, 0xf, 0xf, false));
}

// a double moved within each quad: lane q takes lane PERM's pick of the quad (DPP quad_perm on both
// halves; 0xB1 swaps neighbours, 0x4E swaps pairs)
template <int PERM>
__device__ __forceinline__ double qperm(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), PERM, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), PERM, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// the sum of x over the quad, the same bits on its four lanes ((x0 + x1) + (x2 + x3), each add
// commuted alike)
__device__ __forceinline__ double qsum(double x) {
  const double s = x + qperm<0xB1>(x);
  return s + qperm<0x4E>(s);
}

// The blocked LU of one lane quad's matrix per quad, 16 per wave, in registers: lane q holds column q
// of each 4 x 4 block, block (I, J)'s row i in blk[I][J][i]. Every lane of the wave takes part (the
// MFMAs need every lane).
template <int NB>
__device__ __forceinline__ void lu_factor(v4f (&blk)[NB][NB], int q) {
  sfor<0, NB>([&]<int K>() {
    v4f &D = blk[K][K];
    // the diagonal block: lane q holds its column q; L (unit, strictly below) and U overwrite it
    sfor<0, 4>([&]<int p>() {
      const float rp = 1.0f / qb<p>(D[p]);
      sfor<p + 1, 4>([&]<int i>() {
        const float li = qb<p>(D[i]) * rp;
        D[i] = q > p ? D[i] - li * D[p] : (q == p ? li : D[i]);
      });
    });
    // L_KK[i][p], broadcast to the quad
    float Lkk[4][4];
    sfor<0, 4>([&]<int p>() { sfor<p + 1, 4>([&]<int i>() { Lkk[i][p] = qb<p>(D[i]); }); });
    // U_KJ = L_KK^-1 A_KJ right of the diagonal, column by column in each lane
    sfor<K + 1, NB>([&]<int J>() {
      v4f &X = blk[K][J];
      sfor<0, 3>([&]<int p>() { sfor<p + 1, 4>([&]<int i>() { X[i] = X[i] - Lkk[i][p] * X[p]; }); });
    });
    // L_IK = A_IK U_KK^-1 below it, row by row across the quad; aL[I][p] is the MFMA operand of
    // column p, lane 4b + i holding -L_IK[i][p]
    const float ud = q == 0 ? D[0] : q == 1 ? D[1] : q == 2 ? D[2] : D[3];
    const float rud = 1.0f / ud;
    float aL[NB][4];
    sfor<K + 1, NB>([&]<int I>() {
      v4f &X = blk[I][K];
      sfor<0, 4>([&]<int p>() { aL[I][p] = 0.0f; });
      sfor<0, 4>([&]<int i>() {
        sfor<0, 4>([&]<int p>() {
          const float xp = qb<p>(X[i] * rud);
          X[i] = q > p ? X[i] - xp * D[p] : (q == p ? xp : X[i]);
          aL[I][p] = q == i ? -xp : aL[I][p];
        });
      });
    });
    // the Schur complement on the matrix cores: A_IJ += sum_p (-L_IK[:, p]) U_KJ[p, :]
    sfor<K + 1, NB>([&]<int I>() {
      sfor<K + 1, NB>([&]<int J>() {
        sfor<0, 4>([&]<int p>() {
          blk[I][J] = __builtin_amdgcn_mfma_f32_4x4x1f32(aL[I][p], blk[K][J][p], blk[I][J], 0, 0, 0);
        });
      });
    });
  });
}

// The batched LU of 16 matrices per wave, in place in the group-major array W. A wave exits
// whole; the lanes of a group's missing matrices factor the identity.
template <int NB>
__global__ __launch_bounds__(256) void k_lu(float *__restrict__ W, int64_t ncell) {
  const int64_t grp = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (grp * IM_G >= ncell) return;
  const int lane = threadIdx.x & 63, b = lane >> 2, q = lane & 3;
  const bool real = grp * IM_G + b < ncell;
  float *base = W + grp * IM_GROUP + b;
  v4f blk[NB][NB];
  sfor<0, NB>([&]<int I>() {
    sfor<0, NB>([&]<int J>() {
      sfor<0, 4>([&]<int i>() {
        const int r = 4 * I + i, c = 4 * J + q;
        blk[I][J][i] = real ? base[(r * IM_N + c) * IM_G] : (r == c ? 1.0f : 0.0f);
      });
    });
  });
  lu_factor<NB>(blk, q);
  if (!real) return;
  sfor<0, NB>([&]<int I>() {
    sfor<0, NB>([&]<int J>() {
      sfor<0, 4>([&]<int i>() { base[((4 * I + i) * IM_N + 4 * J + q) * IM_G] = blk[I][J][i]; });
    });
  });
}

// x = (L U)^-1 z in FP32 from the factors of matrix `cell`, in place (L unit lower)
template <int NV>
__device__ __forceinline__ void lu_solve(const float *__restrict__ W, int64_t cell, float (&z)[NV]) {
#pragma unroll
  for (int r = 1; r < NV; r++) {
#pragma unroll
    for (int c = 0; c < r; c++) z[r] -= W[lu_at(cell, r, c)] * z[c];
  }
#pragma unroll
  for (int r = NV - 1; r >= 0; r--) {
#pragma unroll
    for (int c = r + 1; c < NV; c++) z[r] -= W[lu_at(cell, r, c)] * z[c];
    z[r] = z[r] / W[lu_at(cell, r, r)];
  }
}

struct ImArgs {
  const double *K, *C;
  View Q, st, out;
  int begin_dQ, first, last;
  double *A;     // the FP64 matrices, group-major with nv^2 entries
  double *y;     // [ns + 1][ncell] three times: the part's start (then its end), f at its FP32 rounding, f at it
  double *tc;    // [im_cache_n][ncell]: the cache of f at the FP32 rounding's T
  int64_t ncell;
  int32_t *sweeps;
  const double *dt;
  Box3 b;
  int parts, maxsweeps;
  double tol;
};

// entry (r, c) of FP64 matrix `cell` with nv unknowns
template <int NV>
__device__ __forceinline__ int64_t a64_at(int64_t cell, int r, int c) {
  return (cell >> 4) * (NV * NV * IM_G) + (int64_t)(r * NV + c) * IM_G + (cell & 15);
}

// f(y) of the 0D reactor at constant density: the species' rates over rho and dT/dt at constant
// volume, as the sub-stepped form takes them (substep.mako:14-50: ln T from the rate state, cv and
// each species' internal energy by their ln T polynomials)
template <class M>
__device__ __forceinline__ void rhs0d(const Gas<M::NS> &g, const double *__restrict__ C, double rho, double rhoinv,
                                      const double (&y)[M::NS + 1], double (&f)[M::NS + 1]) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS;
  double Y[NS], rate[NS], D[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) Y[n] = y[n];
  const double T = y[NS];
  typename M::Rates s;
  M::ratestate(g, C, rho, Y, T, s);
  M::template fold<false>(C, s, rate, D);
  double cv = 0.0, dTdt = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const double cpR = cpR_species(g, n, s.logT);
    const double e = e_species(g, n, T, s.logT, s.Tinv);
    cv += (cpR - 1.0) * g.Rsp(n) * Y[n];
    dTdt -= e * rate[n];
  }
  const double cvrhoinv = 1.0 / (cv * rho);
#pragma unroll
  for (int n = 0; n < NS; n++) f[n] = rate[n] * rhoinv;
  f[NS] = dTdt * cvrhoinv;
}

// What f(y) shares over every composition at one T: the rates' T-only values (M::TCache: ln T, 1/T,
// the Gibbs terms and their exponentials, each reaction's cached rate constant, ln k or Troe fit) and
// each species' cv share (cp/R - 1) R_n and internal energy e_n. A forward-difference column of a
// species changes the composition at the same T, so it takes f from this and the composition alone;
// every value is the expression rhs0d evaluates, so the column keeps rhs0d's bits.
template <class M>
struct ImCache {
  typename M::TCache t;
  double ea[M::NS], ee[M::NS];
};

template <class M>
constexpr int im_cache_n = sizeof(ImCache<M>) / sizeof(double);

// the cache of f at T into tc[entry * nc] in ImCache's order, each value stored as it is made (C, K and
// tc in one restrict scope, so the constants keep their scalar loads and nothing waits in a register)
template <class M>
__device__ __forceinline__ void im_cache_write(const double *__restrict__ K, const double *__restrict__ C, double T,
                                               double *__restrict__ tc, int64_t nc) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS, NT = sizeof(typename M::TCache) / sizeof(double);
  const Gas<NS> g{K};
  M::tstate(g, C, T, tc, nc);
  const double logT = tlog(T), Tinv = 1.0 / T;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const double cpR = cpR_species(g, n, logT);
    tc[(NT + n) * nc] = (cpR - 1.0) * g.Rsp(n);
    tc[(NT + NS + n) * nc] = e_species(g, n, T, logT, Tinv);
  }
}

// rhs0d at a composition whose T is the cache's
template <class M>
__device__ __forceinline__ void rhs0d_cached(const Gas<M::NS> &g, const double *__restrict__ C, double rho,
                                             double rhoinv, const double (&y)[M::NS + 1], const ImCache<M> &c,
                                             double (&f)[M::NS + 1]) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS;
  double Y[NS], rate[NS], D[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) Y[n] = y[n];
  typename M::Rates s;
  M::cstate(g, C, rho, Y, c.t, s);
  M::template fold_cached<false>(C, s, c.t, rate, D);
  double cv = 0.0, dTdt = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    cv += c.ea[n] * Y[n];
    dTdt -= c.ee[n] * rate[n];
  }
  const double cvrhoinv = 1.0 / (cv * rho);
#pragma unroll
  for (int n = 0; n < NS; n++) f[n] = rate[n] * rhoinv;
  f[NS] = dTdt * cvrhoinv;
}

// the cache of work place `code` from the work array's [entry][ncell] region
template <class M>
__device__ __forceinline__ void im_cache_load(const double *__restrict__ tc, int64_t nc, int64_t code, ImCache<M> &c) {
  double v[im_cache_n<M>];
#pragma unroll
  for (int e = 0; e < im_cache_n<M>; e++) v[e] = tc[e * nc + code];
  __builtin_memcpy(&c, v, sizeof(v));
}

__device__ __forceinline__ int64_t im_code(const Box3 &b, int i, int j, int k) {
  return (i - b.s0) + (int64_t)b.e0 * ((j - b.s1) + (int64_t)b.e1 * (k - b.s2));
}

// the step of unknown m's forward difference: 2^-14 of T or of a mass fraction, at least 2^-14 * 1e-3
// for a trace species. Far above the square root of the rounding unit on purpose: the rates'
// rounding divided by the step is what the difference adds of its own, and a step of 2^-26 would
// turn a last-bit change of y into a 1e-6 change of a stiff part's result and mass drift through
// the columns' sums; the step's truncation is smooth in y, and the linearly implicit Euler step
// is first order with any fixed matrix. The step taken is the difference the rounding leaves
__device__ __forceinline__ double fd_step(double ym, bool temperature) {
#pragma clang fp contract(off)
  const double scale = temperature ? ym : fmax(fabs(ym), 1e-3);
  return (ym + 6.103515625e-05 * scale) - ym;
}

// x = A^-1 b by iterative refinement on the FP32 factors W (the batched solve of caller-given
// systems), the residual b - A x in FP64 against A (entry (r, c) at A[at(r, c)]); a lane per system;
// stops when every row's residual is within tol of its scale
// |b_r| + sum_c |A_rc x_c|, or after maxsweeps corrections. Returns the corrections taken, or
// maxsweeps + 1 when the last one left a row outside the tolerance
template <int NV, class At>
__device__ __forceinline__ int refine(const float *__restrict__ W, const double *__restrict__ A, At at, int64_t cell,
                                      const double (&bv)[NV], double (&x)[NV], int maxsweeps, double tol) {
#pragma clang fp contract(off)
  double r[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    x[m] = 0.0;
    r[m] = bv[m];
  }
  int s = 0;
  bool done = false;
#pragma unroll 1
  while (s < maxsweeps && !done) {
    const float *Ws = opaque(W);
    const double *As = opaque(A);
    float z[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) z[m] = (float)r[m];
    lu_solve<NV>(Ws, cell, z);
#pragma unroll
    for (int m = 0; m < NV; m++) x[m] += (double)z[m];
    done = true;
#pragma unroll
    for (int m = 0; m < NV; m++) {
      double acc = bv[m], scale = fabs(bv[m]);
#pragma unroll
      for (int c = 0; c < NV; c++) {
        const double ax = As[at(cell, m, c)] * x[c];
        acc = acc - ax;
        scale = scale + fabs(ax);
      }
      r[m] = acc;
      done = done && fabs(acc) <= tol * scale;
    }
    s++;
  }
  return done ? s : maxsweeps + 1;
}

// A = I - h J at the part's start, J by forward differences, in three launches: the cache (a lane per
// cell) stores y and what f shares over the compositions of y^'s T (y^: y rounded to FP32); the base
// f(y^); the columns (a lane per cell and unknown) each one column of A, a species column from the
// cache, the T column by two whole evaluations, and one more slice f(y) itself, the solve's
// right-hand side. J is taken at y^: a
// last-bit change of y (the oracle's FP64 solve against the device's refined one, a part's start
// after another part) then leaves the matrix as it was, where the differences' own rounding would
// move it by 1e-12 of itself and a stiff part's result by more; the linearly implicit Euler step is
// first order with the matrix of any nearby state. A lane makes one evaluation of the rates, so the
// lanes hold the registers of one evaluation and the chip's waves are as many as for the rates alone.

// the part's start: the first part takes y from Q and the state, a later one the y the last part left
// (selected rather than branched, one copy of the code); y and the cache of f at y^'s T into the work
// array, every store after the constants' loads (which then go through the scalar cache)
template <class M>
__device__ __forceinline__ void ijac_tcache(const ImArgs &a, const View &Q, const View &st, int i, int j, int k,
                                            int64_t code) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS, NV = NS + 1;
  const int64_t nc = a.ncell;
  const int64_t o = Q.off(i, j, k);
  const double rhoinv = 1.0 / Q.at(o, 0);
  double *y = a.y + code;
  double y0[NV], Y[NS];
  massfrac<NS>(Q, o, rhoinv, Y);
#pragma unroll
  for (int n = 0; n < NS; n++) y0[n] = a.first ? Y[n] : y[n * nc];
  y0[NS] = a.first ? st.at(st.off(i, j, k), 1) : y[NS * nc];
  im_cache_write<M>(a.K, a.C, (double)(float)y0[NS], a.tc + code, nc);
#pragma unroll
  for (int r = 0; r < NV; r++) y[r * nc] = y0[r];
}

// the base: f at y rounded to FP32 (y^), from the cache, into the work array
template <class M>
__device__ __forceinline__ void ijac_base(const ImArgs &a, const View &Q, int i, int j, int k, int64_t code) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS, NV = NS + 1;
  const int64_t nc = a.ncell;
  const double rho = Q.at(Q.off(i, j, k), 0), rhoinv = 1.0 / rho;
  const double *y = a.y + code;
  double yc[NV], fc[NV];
#pragma unroll
  for (int r = 0; r < NV; r++) yc[r] = (double)(float)y[r * nc];
  ImCache<M> cc;
  im_cache_load<M>(a.tc, nc, code, cc);
  rhs0d_cached<M>(Gas<NS>{a.K}, a.C, rho, rhoinv, yc, cc, fc);
  double *fh = a.y + NV * nc + code;
#pragma unroll
  for (int r = 0; r < NV; r++) fh[r * nc] = fc[r];
}

// slice m (wave-uniform) of the columns' launch: for m < ns + 1, column m of A, the rates at
// y^ + dlt e_m less f(y^), over dlt, times -h, plus the identity's column; for m = ns + 1, f(y)
template <class M>
__device__ __forceinline__ void ijac_column(const ImArgs &a, const View &Q, int i, int j, int k, int64_t code,
                                            int m) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS, NV = NS + 1;
  const int64_t nc = a.ncell;
  const int64_t o = Q.off(i, j, k);
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  const double *y = a.y + code, *fh = a.y + NV * nc + code;
  const bool column = m < NV;
  double yp[NV], f1[NV], ym = 0.0;
#pragma unroll
  for (int r = 0; r < NV; r++) {
    const double yr = y[r * nc];
    yp[r] = column ? (double)(float)yr : yr;
    ym = r == m ? yp[r] : ym;
  }
  const double dlt = fd_step(ym, m == NS);
#pragma unroll
  for (int r = 0; r < NV; r++) yp[r] = r == m ? ym + dlt : yp[r];
  if (m < NS) {
    // a species column: T is the base's, whose cache the base left
    ImCache<M> cc;
    im_cache_load<M>(a.tc, nc, code, cc);
    rhs0d_cached<M>(Gas<NS>{a.K}, a.C, rho, rhoinv, yp, cc, f1);
  } else {
    rhs0d<M>(Gas<NS>{a.K}, a.C, rho, rhoinv, yp, f1);
  }
  if (!column) {
    double *bv = a.y + 2 * NV * nc + code;
#pragma unroll
    for (int r = 0; r < NV; r++) bv[r * nc] = f1[r];
    return;
  }
  const double h = *a.dt / a.parts;
  const double hd = h / dlt;
#pragma unroll
  for (int r = 0; r < NV; r++) a.A[a64_at<NV>(code, r, m)] = (r == m ? 1.0 : 0.0) - (f1[r] - fh[r * nc]) * hd;
}

// what a slice of the columns' launch leaves for a work place that holds no cell (a tile's overhang in
// a table of blocks): the identity's column, or a zero right-hand side, which the solve takes in one
// sweep
template <int NV>
__device__ __forceinline__ void ijac_void(const ImArgs &a, int64_t code, int m) {
  if (m < NV) {
#pragma unroll
    for (int r = 0; r < NV; r++) a.A[a64_at<NV>(code, r, m)] = r == m ? 1.0 : 0.0;
  } else {
#pragma unroll
    for (int r = 0; r < NV; r++) a.y[2 * NV * a.ncell + r * a.ncell + code] = 0.0;
  }
}

// The part's solve and step, a lane quad per cell, 16 cells per wave: (I - h J) x = f(y) by iterative
// refinement, then y += h x. The quad loads its FP64 matrix once (lane q: columns q, q + 4, q + 8)
// and keeps it, and the FP32 factors of its rounding (lu_factor), in registers; each sweep solves
// L U z = r in FP32 with the residual broadcast to the quad, adds z to x and takes r = f - A x in
// FP64 (each lane's columns, summed over the quad) with the row scales |f_m| + sum_c |A_mc x_c|. A
// cell stops when every row's residual is within tol of its scale (its x then stays as it is), or
// after maxsweeps; the wave runs until its last cell stops. Lane q keeps rows q, q + 4, q + 8 of x, f
// and r. A cell's x depends only on its own matrix and f, whichever cells share its wave.
template <int NB, int NV>
__device__ __forceinline__ void lu_refine_group(const ImArgs &a, int64_t grp, int lane) {
#pragma clang fp contract(off)
  const int b = lane >> 2, q = lane & 3;
  const int64_t nc = a.ncell, cell = grp * IM_G + b;
  const bool real = cell < nc;
  const int64_t at = real ? cell : grp * IM_G;
  double Ad[NB][NV];
  sfor<0, NB>([&]<int J>() {
    const int c = 4 * J + q;
    const int cc = c < NV ? c : 0;
    sfor<0, NV>([&]<int m>() {
      const double v = a.A[a64_at<NV>(at, m, cc)];
      Ad[J][m] = !real ? (m == c ? 1.0 : 0.0) : (c < NV ? v : 0.0);
    });
  });
  v4f blk[NB][NB];
  sfor<0, NB>([&]<int I>() {
    sfor<0, NB>([&]<int J>() {
      sfor<0, 4>([&]<int i>() {
        constexpr int r = 4 * I + i;
        if constexpr (r < NV)
          blk[I][J][i] = (float)Ad[J][r];
        else
          blk[I][J][i] = r == 4 * J + q ? 1.0f : 0.0f;
      });
    });
  });
  lu_factor<NB>(blk, q);
  // the diagonal's reciprocals, to the quad
  float rdiag[NV];
  sfor<0, NV>([&]<int c>() { rdiag[c] = 1.0f / qb<c % 4>(blk[c / 4][c / 4][c % 4]); });
  const double *bvp = a.y + 2 * NV * nc + at;
  double bv[NB], x[NB], r[NB];
  sfor<0, NB>([&]<int t>() {
    const int m = 4 * t + q;
    const double v = bvp[(m < NV ? m : 0) * nc];
    bv[t] = real && m < NV ? v : 0.0;
    x[t] = 0.0;
    r[t] = bv[t];
  });
  // the flags are ints and a cell's count is set inside the loop: no lane mask lives out of it
  int done = 0, count = a.maxsweeps + IM_FAIL;
  int s = 0;
#pragma unroll 1
  while (s < a.maxsweeps && !__all(done)) {
    // z = (L U)^-1 r in FP32, replicated over the quad, in lu_solve's order
    float z[NV];
    sfor<0, NV>([&]<int m>() { z[m] = qb<m % 4>((float)r[m / 4]); });
    sfor<1, NV>([&]<int m>() { sfor<0, m>([&]<int c>() { z[m] -= qb<c % 4>(blk[m / 4][c / 4][m % 4]) * z[c]; }); });
    sfor<0, NV>([&]<int n>() {
      constexpr int m = NV - 1 - n;
      sfor<m + 1, NV>([&]<int c>() { z[m] -= qb<c % 4>(blk[m / 4][c / 4][m % 4]) * z[c]; });
      z[m] = z[m] * rdiag[m];
    });
    // x += z on the lanes' own rows, for the cells still refining
    sfor<0, NB>([&]<int t>() {
      float zq = 0.0f;
      sfor<0, 4>([&]<int p>() {
        if constexpr (4 * t + p < NV) zq = q == p ? z[4 * t + p] : zq;
      });
      x[t] = done ? x[t] : x[t] + (double)zq;
    });
    // r = f - A x and the rows' scales, the products over the lane's columns summed over the quad
    int ok = 1;
    sfor<0, NV>([&]<int m>() {
      double p = 0.0, sc = 0.0;
      sfor<0, NB>([&]<int J>() {
        const double ax = Ad[J][m] * x[J];
        p = p + ax;
        sc = sc + fabs(ax);
      });
      p = qsum(p);
      sc = qsum(sc);
      if (q == m % 4) {
        const double acc = bv[m / 4] - p, scale = fabs(bv[m / 4]) + sc;
        r[m / 4] = done ? r[m / 4] : acc;
        ok = ok & (fabs(acc) <= a.tol * scale ? 1 : 0);
      }
    });
    // the cell's rows are spread over its quad: done when all four lanes' rows are
    ok = ok & __builtin_amdgcn_mov_dpp(ok, 0xB1, 0xf, 0xf, false);
    ok = ok & __builtin_amdgcn_mov_dpp(ok, 0x4E, 0xf, 0xf, false);
    s++;
    count = !done && ok ? s : count;
    done = done | ok;
  }
  if (!real) return;
  if (q == 0) a.sweeps[cell] = a.first ? count : a.sweeps[cell] + count;
  const double h = *a.dt / a.parts;
  sfor<0, NB>([&]<int t>() {
    const int m = 4 * t + q;
    if (m < NV) a.y[m * nc + cell] = a.y[m * nc + cell] + h * x[t];
  });
}

// the last part's end: the mean rate over the step, rho (Y_end - Y_start) / dt, into out as the
// sub-stepped form begins dQ (or into the rates)
template <class M>
__device__ __forceinline__ void iend_cell(const ImArgs &a, const View &Q, const View &out, int i, int j, int k,
                                          int64_t code) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS;
  const int64_t nc = a.ncell;
  const int64_t o = Q.off(i, j, k);
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  double Ys[NS];
  massfrac<NS>(Q, o, rhoinv, Ys);
  const double rdt = rho / *a.dt;
  const int64_t oo = out.off(i, j, k);
  if (a.begin_dQ) {
#pragma unroll
    for (int l = 0; l < 5; l++) out.at(oo, l) = 0.0;
#pragma unroll
    for (int n = 0; n < NS - 1; n++) out.at(oo, 5 + n) = (a.y[n * nc + code] - Ys[n]) * rdt;
  } else {
#pragma unroll
    for (int n = 0; n < NS - 1; n++) out.at(oo, n) = (a.y[n * nc + code] - Ys[n]) * rdt;
  }
}

template <class M>
__global__ __launch_bounds__(FLAT) void k_itc(ImArgs a) {
  int i, j, k;
  if (!flat_cell_of(a.b, i, j, k)) return;
  ijac_tcache<M>(a, a.Q, a.st, i, j, k, im_code(a.b, i, j, k));
}

template <class M>
__global__ __launch_bounds__(FLAT) void k_ibase(ImArgs a) {
  int i, j, k;
  if (!flat_cell_of(a.b, i, j, k)) return;
  ijac_base<M>(a, a.Q, i, j, k, im_code(a.b, i, j, k));
}

// every slice of a tile of cells (the ns + 1 columns and f(y)) in workgroups numbered one after
// another, so one XCD runs a tile's slices and its L2 serves y and f(y^) to all of them
template <class M>
__global__ __launch_bounds__(FLAT) void k_icol(ImArgs a) {
  constexpr unsigned S = M::NS + 2;
  const unsigned g = xcd_block(), tile = g / S;
  int i, j, k;
  if (!flat_at(a.b, tile, i, j, k)) return;
  ijac_column<M>(a, a.Q, i, j, k, im_code(a.b, i, j, k), (int)(g - tile * S));
}

// four waves of 16 cells per workgroup
template <int NB, int NV>
__global__ __launch_bounds__(256) void k_lusolve(ImArgs a) {
  const int64_t grp = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (grp * IM_G >= a.ncell) return;
  lu_refine_group<NB, NV>(a, grp, threadIdx.x & 63);
}

template <class M>
__global__ __launch_bounds__(FLAT) void k_iend(ImArgs a) {
  int i, j, k;
  if (!flat_cell_of(a.b, i, j, k)) return;
  iend_cell<M>(a, a.Q, a.out, i, j, k, im_code(a.b, i, j, k));
}

}  // namespace pga
