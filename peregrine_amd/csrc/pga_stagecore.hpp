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

// The divergence, the source and the RK combine of the cells (i, j, kbeg..kend-1) of one column:
// memory-bound
template <int NE, bool CHEM>
__device__ __forceinline__ void stage_cells(const StageArgs &a, int i, int j, int kbeg, int kend) {
  const View Qin = a.Qin, Q0 = a.Q0, Qout = a.Qout, det = a.det;
  const View iF = a.iF, jF = a.jF, kF = a.kF;
  const CombineSpec comb = a.comb;
  const View rates = a.rates;
  const double dt = *a.dt;
#pragma unroll 1
  for (int k = kbeg; k < kend; k++) {
    const int64_t o = Qin.off(i, j, k), o0 = Q0.off(i, j, k), oo = Qout.off(i, j, k);
    const int64_t fil = iF.off(i, j, k), fjl = jF.off(i, j, k), fkl = kF.off(i, j, k);
    const int64_t fih = fil + 1, fjh = fjl + jF.sj, fkh = fkl + kF.sk;
    const int64_t orr = CHEM ? rates.off(i, j, k) : 0;
    const double J = __builtin_nontemporal_load(&det.at(det.off(i, j, k), 0));
    // per conserved row: its six faces in the reference's order of sums, times detJinv, the
    // source first where there is one (chemint/kernels/source.mako:3-10), then the RK combine in
    // the reference's order of terms (integrators/kernels/combine.mako:1-13); a row at a time in a
    // rolled loop keeps ~30 registers, so eight waves per SIMD hide the loads' latency. What only
    // this cell reads (Q, Q0, the rates, 1/J) and what it writes stream past the caches
    // (nontemporal); each face value is read by the two cells it lies between and is cached
#pragma unroll 1
    for (int l = 0; l < NE; l++) {
#pragma clang fp contract(off)
      double d = ((((iF.at(fil, l) + jF.at(fjl, l)) + kF.at(fkl, l)) - iF.at(fih, l)) - jF.at(fjh, l)) - kF.at(fkh, l);
      d = d * J;
      if constexpr (CHEM) {
        if (l >= 5) d = __builtin_nontemporal_load(&rates.at(orr, l - 5)) + d;
      }
      // the terms' values are selected as values: a select of pointers would put them in memory
      const double qin = __builtin_nontemporal_load(&Qin.at(o, l));
      const double q0 = a.q0 ? __builtin_nontemporal_load(&Q0.at(o0, l)) : 0.0;
      double va = 0.0;
      if (a.two) {
        // the first combine, its own register read as it stood, written before the second reads it
        const int64_t oa = a.QA.off(i, j, k);
        const double own = a.QA.at(oa, l);
#pragma unroll
        for (int t = 0; t < 4; t++) {
          if (t < a.combA.n) {
            const int src = a.combA.src[t];
            const double x = src == SRC_QIN ? qin : src == SRC_Q0 ? q0 : src == SRC_OWNA ? own : d;
            const double term = a.combA.withdt[t] ? a.combA.w[t] * dt * x : a.combA.w[t] * x;
            va = t == 0 ? term : va + term;
          }
        }
        a.QA.at(oa, l) = va;
      }
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if (t < comb.n) {
          const int src = comb.src[t];
          const double x = src == SRC_QIN ? qin : src == SRC_Q0 ? q0 : src == SRC_A ? va : d;
          const double term = comb.withdt[t] ? comb.w[t] * dt * x : comb.w[t] * x;
          acc = t == 0 ? term : acc + term;
        }
      }
      __builtin_nontemporal_store(acc, &Qout.at(oo, l));
    }
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
  const int64_t od = Qdst.off(i, j, k);
  double Q[NS + 4];
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Q[l] = Qsrc.at(os, l);
  if (turn.on) {
    const double x = Q[1], y = Q[2], z = Q[3];
    turned(turn.R, x, y, z, Q[1], Q[2], Q[3]);
  }
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Qdst.at(od, l) = Q[l];
  const Gas<NS> g{K};
  const int64_t ost = stdst.off(i, j, k);
  EosState s;
  fromQ_regs<NS, EOS>(g, Q, stguess.at(ost, 1), s);
  store_state<EOS>(stdst, ost, s);
}

}  // namespace pga
