// The per-cell work of the fused stage, shared by the block-at-a-time launches (pga_stage.hip) and
// the launches over a table of blocks (pga_batch.hip), so both run the same arithmetic to the bit:
// the divergence + source + RK combine of a column of cells, the EOS of a cell's new Q, a cell's
// production rates, and a halo cell's copy and EOS.
#pragma once

#include "generated/mechanisms.hpp"
#include "pga_chem.hpp"
#include "pga_flux.hpp"

namespace pga {


// a term's source: the stage's input, the step's start, the stage's derivative, the first combine's
// output register as it stood (its own term), the first combine's value just taken
enum { SRC_QIN = 0, SRC_Q0 = 1, SRC_DQ = 2, SRC_OWNA = 3, SRC_A = 4 };

// out = sum_t w_t x_t in order, a term with withdt taking w_t * dt * x_t
struct CombineSpec {
  int n;
  int src[4];
  int withdt[4];
  double w[4];
};

struct StageArgs {
  const double *K;
  View Qin, stin, Q0, Qout, stout, det, rates, iF, jF, kF;
  CombineSpec comb;
  // a stage with two combines (rk4: the running total, then the next stage or the step's end): the
  // first, combA, into QA before comb into Qout
  View QA;
  CombineSpec combA;
  int two;
  const double *dt;
  int s0, s1, s2, n0, n1, n2;
  int kc;
  int q0;  // whether a combine term reads Q0 (RK3's first stage reads only Qin and dQ)
};

constexpr int WI = 64, WJ = 4, WT = WI * WJ;

// the EOS inversion of a cell's new Q held in registers, from T guess
template <int NS, int EOS>
__device__ __forceinline__ void fromQ_regs(const Gas<NS> &g, const double (&Q)[NS + 4], double Tguess, EosState &s) {
#pragma clang fp contract(off)
  const double rho = Q[0], rhoinv = 1.0 / rho;
  const double e = (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * rhoinv) * rhoinv;
  double Y[NS > 1 ? NS - 1 : 1];
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Y[n] = Q[5 + n] * rhoinv;
  MixMom<NS, EOS> m;
  mixmom<NS, EOS>(g, [&](int n) { return Y[n]; }, m);
  fromrhoE<NS, EOS>(g, m, [&](int n) { return Y[n]; }, rho, e, Tguess, s);
}

// The RK combines of row l from its derivative d (the faces' sum times detJinv, the source added):
// the first combine into va (its own register's value `own` as it stood) where the stage has two,
// then the combine returned, each in the reference's order of terms
// (integrators/kernels/combine.mako:1-13)
__device__ __forceinline__ double row_combine(const StageArgs &a, double d, double qin, double q0, double own,
                                              double dt, double &va) {
#pragma clang fp contract(off)
  va = 0.0;
  if (a.two) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (t < a.combA.n) {
        const int src = a.combA.src[t];
        const double x = src == SRC_QIN ? qin : src == SRC_Q0 ? q0 : src == SRC_OWNA ? own : d;
        const double term = a.combA.withdt[t] ? a.combA.w[t] * dt * x : a.combA.w[t] * x;
        va = t == 0 ? term : va + term;
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    if (t < a.comb.n) {
      const int src = a.comb.src[t];
      const double x = src == SRC_QIN ? qin : src == SRC_Q0 ? q0 : src == SRC_A ? va : d;
      const double term = a.comb.withdt[t] ? a.comb.w[t] * dt * x : a.comb.w[t] * x;
      acc = t == 0 ? term : acc + term;
    }
  }
  return acc;
}

// Row l of one cell's stage: its six faces in the reference's order of sums (the k faces given), times
// detJinv, the source first where there is one (chemint/kernels/source.mako:3-10), then the RK combine
// in the reference's order of terms (integrators/kernels/combine.mako:1-13). What only this cell reads
// (Q, Q0, the rates, 1/J) and what it writes stream past the caches (nontemporal); each face value is
// read by the two cells it lies between and is cached
template <bool CHEM>
__device__ __forceinline__ void stage_row(const StageArgs &a, int l, double kfl, double kfh, int64_t o, int64_t o0,
                                          int64_t oo, int64_t fil, int64_t fjl, int64_t orr, double J, double dt,
                                          int i, int j, int k) {
#pragma clang fp contract(off)
  const View &iF = a.iF, &jF = a.jF;
  const int64_t fih = fil + 1, fjh = fjl + jF.sj;
  double d = ((((iF.at(fil, l) + jF.at(fjl, l)) + kfl) - iF.at(fih, l)) - jF.at(fjh, l)) - kfh;
  d = d * J;
  if constexpr (CHEM) {
    if (l >= 5) d = __builtin_nontemporal_load(&a.rates.at(orr, l - 5)) + d;
  }
  // the terms' values are selected as values: a select of pointers would put them in memory
  const double qin = __builtin_nontemporal_load(&a.Qin.at(o, l));
  const double q0 = a.q0 ? __builtin_nontemporal_load(&a.Q0.at(o0, l)) : 0.0;
  double va;
  if (a.two) {
    // the first combine, its own register read as it stood, written before the second reads it
    const int64_t oa = a.QA.off(i, j, k);
    const double acc = row_combine(a, d, qin, q0, a.QA.at(oa, l), dt, va);
    a.QA.at(oa, l) = va;
    __builtin_nontemporal_store(acc, &a.Qout.at(oo, l));
    return;
  }
  __builtin_nontemporal_store(row_combine(a, d, qin, q0, 0.0, dt, va), &a.Qout.at(oo, l));
}

// One cell's stage with everything it reads loaded before anything is written, its new Q returned in
// Qn: for a few rows (NE <= 8), where the rows' loads fit registers, so a thread waits on memory once
// a cell rather than once a row. stage_row's arithmetic to the bit; the caller also loads stin's T
// with the rest
template <int NE, bool CHEM>
__device__ __forceinline__ void stage_cell_regs(const StageArgs &a, int i, int j, int k, double dt,
                                                double (&Qn)[NE]) {
#pragma clang fp contract(off)
  static_assert(NE <= 8, "stage_cell_regs holds every row of a cell in registers");
  const View &iF = a.iF, &jF = a.jF, &kF = a.kF;
  const int64_t o = a.Qin.off(i, j, k), o0 = a.Q0.off(i, j, k), oo = a.Qout.off(i, j, k);
  const int64_t fil = iF.off(i, j, k), fjl = jF.off(i, j, k), fkl = kF.off(i, j, k);
  const int64_t fih = fil + 1, fjh = fjl + jF.sj, fkh = fkl + kF.sk;
  const int64_t orr = CHEM ? a.rates.off(i, j, k) : 0, oa = a.two ? a.QA.off(i, j, k) : 0;
  const double J = __builtin_nontemporal_load(&a.det.at(a.det.off(i, j, k), 0));
  double d[NE], qin[NE], q0[NE], own[NE];
#pragma unroll
  for (int l = 0; l < NE; l++) {
    d[l] = ((((iF.at(fil, l) + jF.at(fjl, l)) + kF.at(fkl, l)) - iF.at(fih, l)) - jF.at(fjh, l)) - kF.at(fkh, l);
    qin[l] = __builtin_nontemporal_load(&a.Qin.at(o, l));
    q0[l] = a.q0 ? __builtin_nontemporal_load(&a.Q0.at(o0, l)) : 0.0;
    own[l] = a.two ? a.QA.at(oa, l) : 0.0;
  }
  double rate[NE];
#pragma unroll
  for (int l = 0; l < NE; l++) rate[l] = CHEM && l >= 5 ? __builtin_nontemporal_load(&a.rates.at(orr, l - 5)) : 0.0;
#pragma unroll
  for (int l = 0; l < NE; l++) {
    double dl = d[l] * J;
    if constexpr (CHEM) {
      if (l >= 5) dl = rate[l] + dl;
    }
    double va;
    Qn[l] = row_combine(a, dl, qin[l], q0[l], own[l], dt, va);
    if (a.two) a.QA.at(oa, l) = va;
    __builtin_nontemporal_store(Qn[l], &a.Qout.at(oo, l));
  }
}

// The divergence, the source and the RK combine of the cells (i, j, kbeg..kend-1) of one column:
// memory-bound. A row at a time in a rolled loop keeps ~30 registers, so eight waves per SIMD hide
// the loads' latency
template <int NE, bool CHEM>
__device__ __forceinline__ void stage_cells(const StageArgs &a, int i, int j, int kbeg, int kend) {
  const View &kF = a.kF;
  const double dt = *a.dt;
#pragma unroll 1
  for (int k = kbeg; k < kend; k++) {
    const int64_t o = a.Qin.off(i, j, k), o0 = a.Q0.off(i, j, k), oo = a.Qout.off(i, j, k);
    const int64_t fil = a.iF.off(i, j, k), fjl = a.jF.off(i, j, k), fkl = kF.off(i, j, k);
    const int64_t fkh = fkl + kF.sk;
    const int64_t orr = CHEM ? a.rates.off(i, j, k) : 0;
    const double J = __builtin_nontemporal_load(&a.det.at(a.det.off(i, j, k), 0));
#pragma unroll 1
    for (int l = 0; l < NE; l++)
      stage_row<CHEM>(a, l, kF.at(fkl, l), kF.at(fkh, l), o, o0, oo, fil, fjl, orr, J, dt, i, j, k);
  }
}

// The k sweep and the stage in one pass (the march): the thread of column (i, j) takes the k faces of
// its chunk one after another, each face's flux computed once into the workgroup's LDS (lo and hi
// slots, [slot][row][thread]) and read by the two cells it lies between, so no k-face array goes to
// HBM; the i and j faces come from their sweeps' arrays. The cells' arithmetic is stage_cells' to the
// bit. The output must not be the input: a chunk's last face reads cells the next chunk writes
template <int NS, int EOS, int RECON, int PRIMARY, int LIM, bool CHEM>
__device__ __forceinline__ void march_cells(const StageArgs &a, const View &kS, double *lds, int i, int j, int kbeg,
                                            int kend) {
  constexpr int NE = NS + 4;
  const int tid = threadIdx.x;
  const double dt = *a.dt;
  int lo = 0;
  {
    double Fv[NE];
    flux_at<NS, EOS, RECON, PRIMARY, LIM>(a.K, 2, a.Qin, a.stin, kS, i, j, kbeg, Fv);
#pragma unroll
    for (int l = 0; l < NE; l++) lds[(lo * NE + l) * WT + tid] = Fv[l];
  }
#pragma unroll 1
  for (int k = kbeg; k < kend; k++) {
    const int hi = lo ^ 1;
    {
      double Fv[NE];
      flux_at<NS, EOS, RECON, PRIMARY, LIM>(opaque(a.K), 2, a.Qin, a.stin, kS, i, j, k + 1, Fv);
#pragma unroll
      for (int l = 0; l < NE; l++) lds[(hi * NE + l) * WT + tid] = Fv[l];
    }
    const int64_t o = a.Qin.off(i, j, k), o0 = a.Q0.off(i, j, k), oo = a.Qout.off(i, j, k);
    const int64_t fil = a.iF.off(i, j, k), fjl = a.jF.off(i, j, k);
    const int64_t orr = CHEM ? a.rates.off(i, j, k) : 0;
    const double J = __builtin_nontemporal_load(&a.det.at(a.det.off(i, j, k), 0));
#pragma unroll 1
    for (int l = 0; l < NE; l++)
      stage_row<CHEM>(a, l, lds[(lo * NE + l) * WT + tid], lds[(hi * NE + l) * WT + tid], o, o0, oo, fil, fjl, orr, J,
                      dt, i, j, k);
    lo = hi;
  }
}

// the EOS of a cell's stage result, T started from the input state (stin may be stout)
template <int NS, int EOS>
__device__ __forceinline__ void eos_cell(const double *__restrict__ K, const View &Q, const View &stin,
                                         const View &stout, int i, int j, int k) {
  const Gas<NS> g{K};
  const int64_t o = Q.off(i, j, k);
  double Qn[NS + 4];
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Qn[l] = Q.at(o, l);
  EosState s;
  fromQ_regs<NS, EOS>(g, Qn, stin.at(stin.off(i, j, k), 1), s);
  store_state<EOS>(stout, stout.off(i, j, k), s);
}

// the carried species' production rates of a cell (chemint/kernels/cell.mako:5-20), the source a
// reacting stage begins dQ with
template <class M>
__device__ __forceinline__ void rates_cell(const double *__restrict__ K, const double *__restrict__ C, const View &Q,
                                           const View &st, const View &rates, int i, int j, int k) {
  constexpr int NS = M::NS;
  const int64_t o = Q.off(i, j, k);
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  double Y[NS], omega[NS];
  massfrac<NS>(Q, o, rhoinv, Y);
  M::production(Gas<NS>{K}, C, rho, Y, st.at(o, 1), omega);
  const int64_t orr = rates.off(i, j, k);
#pragma unroll
  for (int n = 0; n < NS - 1; n++) rates.at(orr, n) = omega[n];
}

// a halo cell: Q copied from the source element os (the momentum turned when `turn` is on) into the
// destination cell (i, j, k), then its EOS into stdst, T started from stguess's
template <int NS, int EOS>
__device__ __forceinline__ void halo_cell(const double *__restrict__ K, const View &Qdst, const View &Qsrc, int64_t os,
                                          const Turn &turn, const View &stdst, const View &stguess, int i, int j,
                                          int k) {
  const int64_t od = Qdst.off(i, j, k), ost = stdst.off(i, j, k);
  double Q[NS + 4];
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Q[l] = Qsrc.at(os, l);
  // the T guess read with Q, before the stores, so the thread waits on memory once
  const double Tguess = stguess.at(ost, 1);
  if (turn.on) {
    const double x = Q[1], y = Q[2], z = Q[3];
    turned(turn.R, x, y, z, Q[1], Q[2], Q[3]);
  }
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Qdst.at(od, l) = Q[l];
  const Gas<NS> g{K};
  EosState s;
  fromQ_regs<NS, EOS>(g, Q, Tguess, s);
  store_state<EOS>(stdst, ost, s);
}

}  // namespace pga
