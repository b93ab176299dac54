// The fused RK stage: after a stage's three flux sweeps (pga_flux, one face per lane), one pass
// over the block's interior does what the reference does in three more launches per stage
// (divergence, combine, interior EOS; simulator/advec/simulator.py:57-87,
// integrators/rungekutta.py:33-39) with the chemistry source added on the way; and the halo pass
// before the sweeps (exchange of Q and the EOS over the halo slabs) is one launch.
//
// A reacting case takes its source from pga_rates, launched before the stage: the production
// rates need ~100 registers of their own.
//
// Each thread owns one (i, j) column of cells and walks a chunk of k-planes. Per cell it sums its
// six face fluxes in the reference's order, adds the chemistry source, takes the RK combine,
// inverts the EOS of the result from the cell's last T, and writes the new Q and state: the face
// fluxes, Q, the step's start and the state are read once and the result written once, with no
// derivative register in between. Output may be the input (each cell reads only itself).
//
// Every cell result is the per-kernel path's to the bit: the same device functions in the same
// order.
#include <type_traits>

#include "generated/mechanisms.hpp"
#include "pga_chem.hpp"
#include "pga_flux.hpp"
#include "pga_internal.hpp"

namespace pga {

enum { SRC_QIN = 0, SRC_Q0 = 1, SRC_DQ = 2 };

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

// the divergence, the source and the RK combine of a column of cells: memory-bound
template <int NE, bool CHEM>
__global__ __launch_bounds__(WT) void k_stage(StageArgs a) {
  const View Qin = a.Qin, Q0 = a.Q0, Qout = a.Qout, det = a.det;
  const View iF = a.iF, jF = a.jF, kF = a.kF;
  const CombineSpec comb = a.comb;
  const View rates = a.rates;
  const int i = a.s0 + blockIdx.x * WI + (threadIdx.x & (WI - 1));
  const int j = a.s1 + blockIdx.y * WJ + threadIdx.x / WI;
  if (i >= a.s0 + a.n0 || j >= a.s1 + a.n1) return;
  const int kbeg = a.s2 + blockIdx.z * a.kc;
  const int kend = min(kbeg + a.kc, a.s2 + a.n2);
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
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if (t < comb.n) {
          const int src = comb.src[t];
          const double x = src == SRC_QIN ? qin : src == SRC_Q0 ? q0 : d;
          const double term = comb.withdt[t] ? comb.w[t] * dt * x : comb.w[t] * x;
          acc = t == 0 ? term : acc + term;
        }
      }
      __builtin_nontemporal_store(acc, &Qout.at(oo, l));
    }
  }
}

// the EOS of the stage's result, T started from the input state (stin may be stout)
template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_stage_eos(const double *__restrict__ K, View Q, View stin, View stout, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  const Gas<NS> g{K};
  const int64_t o = Q.off(i, j, k);
  double Qn[NS + 4];
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Qn[l] = Q.at(o, l);
  EosState s;
  fromQ_regs<NS, EOS>(g, Qn, stin.at(stin.off(i, j, k), 1), s);
  store_state<EOS>(stout, stout.off(i, j, k), s);
}

// the carried species' production rates of each cell (chemint/kernels/cell.mako:5-20), the
// source a reacting stage begins dQ with
template <class M>
__global__ __launch_bounds__(256) void k_rates(const double *__restrict__ K, const double *__restrict__ C, View Q,
                                               View st, View rates, Box3 b) {
  constexpr int NS = M::NS;
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  const int64_t o = Q.off(i, j, k);
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  double Y[NS], omega[NS];
  massfrac<NS>(Q, o, rhoinv, Y);
  M::production(Gas<NS>{K}, C, rho, Y, st.at(o, 1), omega);
  const int64_t orr = rates.off(i, j, k);
#pragma unroll
  for (int n = 0; n < NS - 1; n++) rates.at(orr, n) = omega[n];
}

// the halo pass: Q of each interface's halo slab from its partner's interior, then the EOS
// over the slab (its T guess from the state of the previous stage's halo)
struct HaloTask {
  View Qdst, Qsrc;
  int d0, d1, d2, e0, e1;
  int64_t sbase, st0, st1, st2;  // the source cell of the slab's first cell, the steps of its axes
  int64_t first, cells;          // this slab's first flattened cell index and its cell count
  Turn turn;                     // a rotational periodic's turn of the momentum
};

struct HaloArgs {
  const double *K;
  View stdst, stguess;
  int ntask;
  HaloTask t[6];
};

// one thread per halo cell over the slabs back to back, i fastest within a slab
template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_halo(HaloArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int n = 0;
  while (n + 1 < a.ntask && c >= a.t[n + 1].first) n++;
  const HaloTask &t = a.t[n];
  const int64_t r = c - t.first;
  if (r >= t.cells) return;
  const int ii = (int)(r % t.e0), jj = (int)((r / t.e0) % t.e1), kk = (int)(r / ((int64_t)t.e0 * t.e1));
  const int64_t od = t.Qdst.off(t.d0 + ii, t.d1 + jj, t.d2 + kk), os = t.sbase + ii * t.st0 + jj * t.st1 + kk * t.st2;
  double Q[NS + 4];
#pragma unroll
  for (int l = 0; l < NS + 4; l++) Q[l] = t.Qsrc.at(os, l);
  if (t.turn.on) {
    const double x = Q[1], y = Q[2], z = Q[3];
    turned(t.turn.R, x, y, z, Q[1], Q[2], Q[3]);
  }
#pragma unroll
  for (int l = 0; l < NS + 4; l++) t.Qdst.at(od, l) = Q[l];
  const Gas<NS> g{a.K};
  const int64_t ost = a.stdst.off(t.d0 + ii, t.d1 + jj, t.d2 + kk);
  EosState s;
  fromQ_regs<NS, EOS>(g, Q, a.stguess.at(ost, 1), s);
  store_state<EOS>(a.stdst, ost, s);
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_stage_supported(const pga_case *c) {
  if (!c) return 0;
  int ok = 0;
#define X(N) \
  if (c->desc.ns == N) ok = c->desc.eos != 2 || realgas_compiled(N);
  PGA_NS_CASES(X)
#undef X
  return ok;
}

extern "C" int32_t pga_stage(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view Qout,
                             pga_view stout, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv, pga_view rates,
                             int32_t nterms, const int32_t *src, const double *w, const int32_t *withdt,
                             const double *dt, pga_box box, int32_t kchunk, void *stream) {
  if (!c || !src || !w || !withdt || !dt || nterms < 1 || nterms > 4 || kchunk < 1 || !box_ok(box)) {
    set_error("pga_stage: null argument, 1 to 4 combine terms, a k-chunk of at least 1");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int ne = d.ns + 4;
  const bool chem = d.mechanism >= 0;
  {
    PGA_REQUIRE(covers(Qin, box, ne) && covers(stin, box, nstate(d)) && covers(Q0, box, ne) && covers(Qout, box, ne) &&
                    covers(stout, box, nstate(d)) && covers(detJinv, box, 1),
                "pga_stage: the box reaches outside Qin, Q0, Qout, a state or detJinv");
    const int z[3] = {0, 0, 0}, fh[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    PGA_REQUIRE(covers(iF, box, z, fh[0], ne) && covers(jF, box, z, fh[1], ne) && covers(kF, box, z, fh[2], ne),
                "pga_stage: the box's faces reach outside the face flux arrays");
    for (int t = 0; t < nterms; t++)
      PGA_REQUIRE(src[t] >= 0 && src[t] <= 2, "pga_stage: a combine term reads Qin (0), Q0 (1) or dQ (2)");
    if (chem)
      PGA_REQUIRE(covers(rates, box, d.ns - 1), "pga_stage: a reacting case needs the rates of every cell (pga_rates)");
  }
  StageArgs a{};
  a.K = c->K;
  a.rates = view(rates);
  a.Qin = view(Qin), a.stin = view(stin), a.Q0 = view(Q0), a.Qout = view(Qout), a.stout = view(stout);
  a.iF = view(iF), a.jF = view(jF), a.kF = view(kF), a.det = view(detJinv);
  a.comb.n = nterms;
  for (int t = 0; t < nterms; t++) {
    a.comb.src[t] = src[t], a.comb.w[t] = w[t], a.comb.withdt[t] = withdt[t];
    a.q0 |= src[t] == SRC_Q0;
  }
  a.dt = dt;
  a.s0 = box.start[0], a.s1 = box.start[1], a.s2 = box.start[2];
  a.n0 = box.extent[0], a.n1 = box.extent[1], a.n2 = box.extent[2];
  a.kc = kchunk;
  const dim3 grid((a.n0 + WI - 1) / WI, (a.n1 + WJ - 1) / WJ, (a.n2 + kchunk - 1) / kchunk);
  const Box3 b = box3(box);
  const bool done = dispatch_eos(d, [&]<int N, int E>() {
    if (chem)
      hipLaunchKernelGGL((k_stage<N + 4, true>), grid, dim3(WT), 0, stream_of(stream), a);
    else
      hipLaunchKernelGGL((k_stage<N + 4, false>), grid, dim3(WT), 0, stream_of(stream), a);
    hipLaunchKernelGGL((k_stage_eos<N, E>), grid_of(b), dim3(TI, TJ), 0, stream_of(stream), c->K, a.Qout, a.stin,
                       a.stout, b);
  });
  if (!done) {
    set_error("pga_stage: no kernel for %d species with eos %d", d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_stage");
}

extern "C" int32_t pga_rates(const pga_case *c, pga_view Q, pga_view state, pga_view rates, pga_box box,
                             void *stream) {
  if (!c || !box_ok(box)) {
    set_error("pga_rates: null case or bad box");
    return PGA_EINVAL;
  }
  if (c->desc.mechanism < 0 || !c->chem) {
    set_error("pga_rates: the case has no compiled mechanism");
    return PGA_EUNSUP;
  }
  if (empty(box)) return PGA_OK;
  PGA_REQUIRE(covers(Q, box, c->desc.ns + 4) && covers(state, box, nstate(c->desc)) && covers(rates, box, c->desc.ns - 1),
              "pga_rates: the box reaches outside Q, state or rates");
  const Box3 b = box3(box);
  bool done = false;
#define X(NAME, IDX)                                                                                          \
  if (c->desc.mechanism == IDX) {                                                                             \
    hipLaunchKernelGGL((k_rates<Mech_##NAME>), grid_of(b), dim3(TI, TJ), 0, stream_of(stream), c->K, c->chem, \
                       view(Q), view(state), view(rates), b);                                                 \
    done = true;                                                                                              \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_rates: no mechanism %d", c->desc.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_rates");
}

extern "C" int32_t pga_halo(const pga_case *c, int32_t ntask, const pga_view *Qdst, const pga_view *Qsrc,
                            const int32_t *dst_start, const int32_t *src_start, const int32_t *extent,
                            const int32_t *orient, const double *rot, pga_view stdst, pga_view stguess,
                            void *stream) {
  if (!c || ntask < 0 || ntask > 6 || (ntask && (!Qdst || !Qsrc || !dst_start || !src_start || !extent))) {
    set_error("pga_halo: null argument or more than 6 slabs");
    return PGA_EINVAL;
  }
  const int ne = c->desc.ns + 4;
  HaloArgs a{};
  a.K = c->K;
  a.stdst = view(stdst);
  a.stguess = view(stguess);
  int64_t total = 0;
  for (int n = 0; n < ntask; n++) {
    const int *ds = dst_start + 3 * n, *ss = src_start + 3 * n, *e = extent + 3 * n;
    const pga_box bd{{ds[0], ds[1], ds[2]}, {e[0], e[1], e[2]}};
    PGA_REQUIRE(box_ok(bd) && covers(Qdst[n], bd, ne) && covers(stdst, bd, nstate(c->desc)) &&
                    covers(stguess, bd, nstate(c->desc)),
                "pga_halo: slab %d reaches outside its arrays", n);
    PGA_REQUIRE(Qdst[n].sj == stdst.sj && Qdst[n].sk == stdst.sk, "pga_halo: Q and state must share their layout");
    for (int m = 0; m < 3; m++)
      PGA_REQUIRE(ss[m] >= 0 && e[m] >= 0, "pga_halo: slab %d has a negative source start or extent", n);
    Placement pl;
    if (!placement("pga_halo", Qsrc[n], ss, orient ? orient + 3 * n : nullptr, e, ne, pl)) return PGA_ERANGE;
    HaloTask &t = a.t[n];
    t.Qdst = view(Qdst[n]), t.Qsrc = view(Qsrc[n]);
    t.d0 = ds[0], t.d1 = ds[1], t.d2 = ds[2];
    t.sbase = pl.base, t.st0 = pl.step[0], t.st1 = pl.step[1], t.st2 = pl.step[2];
    t.turn = turn_of(rot ? rot + 9 * n : nullptr);
    t.e0 = e[0], t.e1 = e[1];
    t.first = total;
    t.cells = (int64_t)e[0] * e[1] * e[2];
    total += t.cells;
  }
  a.ntask = ntask;
  if (!total) return PGA_OK;
  const int64_t nblk = (total + 255) / 256;
  PGA_REQUIRE(nblk <= 0x7fffffff, "pga_halo: too many halo cells");
  const dim3 grid((unsigned)nblk);
  const bool done = dispatch_eos(c->desc, [&]<int N, int E>() {
    hipLaunchKernelGGL((k_halo<N, E>), grid, dim3(256), 0, stream_of(stream), a);
  });
  if (!done) {
    set_error("pga_halo: no kernel for %d species with eos %d", c->desc.ns, c->desc.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_halo");
}
