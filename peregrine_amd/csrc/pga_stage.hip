// The fused RK stage: one pass over a block's interior that does what the reference does in
// five launches per stage (source, three flux sweeps, divergence, combine, interior EOS;
// simulator/advec/simulator.py:57-87, integrators/rungekutta.py:33-39), and the halo pass
// before it (exchange of Q and the EOS over the halo slabs).
//
// A reacting case takes its source from pga_rates, launched before the stage: the production
// rates need ~100 registers of their own, and on top of the faces' ~400 they would spill.
//
// Work decomposition. A workgroup of 256 threads owns a 64 (i) x 4 (j) tile of cells and marches
// up a chunk of k-planes. Per plane:
//   - the 65 x 4 i-faces and 64 x 5 j-faces of the tile are computed once each, by the flattened
//     face index across the workgroup, into LDS (60 KB, two workgroups per CU);
//   - each thread computes the k-face above its cell; the one below is the previous plane's,
//     carried in registers;
//   - after a barrier each thread sums its six faces in the reference's order, adds the
//     chemistry source of its cell, takes the RK combine and inverts the EOS of the result,
//     and writes the new Q and state.
// Only a tile's edge faces are taken twice (4 i-faces and 64 j-faces per plane: 9% more face
// work than one per face) and the first k-face of each chunk. The stage reads Q and the state of
// its input once per cell through the caches, and writes Q and the state of its output once.
//
// Every face and cell result is the per-kernel path's to the bit: the same device functions in
// the same order.
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
  View Qin, stin, Q0, Qout, stout, S[3], det, rates;
  CombineSpec comb;
  const double *dt;
  int s0, s1, s2, n0, n1, n2;
  int kc;
};

constexpr int WI = 64, WJ = 4, WT = WI * WJ;
constexpr int NFI = (WI + 1) * WJ;
constexpr int NFJ = WI * (WJ + 1);

// occupancy: the 9-species faces hold ~400 values live, so they run at one wave per SIMD with the
// accumulation registers as spill room; smaller mixtures at two
template <int NS>
struct StageWaves {
  static constexpr int value = NS > 4 ? 1 : 2;
};

// the kernel's arrays, copied out of the argument block so no address of the block is taken
// (which would move it to private memory)
struct StageViews {
  View Qin, stin, S0, S1, S2;
};

template <int NS, int EOS, int RECON, int PRIMARY, int LIM, int AXIS>
__device__ __forceinline__ void face_at(const Gas<NS> g, const StageViews v, int fi, int fj, int fk,
                                        double (&F)[NS + 4]) {
  const int64_t o = v.Qin.off(fi, fj, fk);
  const int64_t d = AXIS == 0 ? 1 : AXIS == 1 ? v.Qin.sj : v.Qin.sk;
  FaceState<NS> L, R;
  face_sides<NS, EOS, RECON, LIM>(g, v.Qin, v.stin, o, d, L, R);
  const View S = AXIS == 0 ? v.S0 : AXIS == 1 ? v.S1 : v.S2;
  const int64_t of = S.off(fi, fj, fk);
  const double Sv[3] = {S.at(of, 0), S.at(of, 1), S.at(of, 2)};
  flux_function<NS, PRIMARY>(L, R, Sv, F);
}

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
  fromrhoE<NS, EOS>(m, rho, e, Tguess, s);
}

template <int NS, int EOS, int RECON, int PRIMARY, int LIM, bool CHEM>
__global__ __launch_bounds__(WT, StageWaves<NS>::value) void k_stage(StageArgs a) {
  constexpr int NE = NS + 4;
  __shared__ double Fi[NE][NFI];
  __shared__ double Fj[NE][NFJ];
  const Gas<NS> g{a.K};
  const StageViews sv{a.Qin, a.stin, a.S[0], a.S[1], a.S[2]};
  const View Qin = a.Qin, stin = a.stin, Q0 = a.Q0, Qout = a.Qout, stout = a.stout, det = a.det;
  const CombineSpec comb = a.comb;
  const View rates = a.rates;
  const int tid = threadIdx.x, lane = tid & (WI - 1), row = tid / WI;
  const int i0 = a.s0 + blockIdx.x * WI, j0 = a.s1 + blockIdx.y * WJ;
  const int iend = a.s0 + a.n0, jend = a.s1 + a.n1;
  const int i = i0 + lane, j = j0 + row;
  const bool active = i < iend && j < jend;
  const int kbeg = a.s2 + blockIdx.z * a.kc;
  const int kend = min(kbeg + a.kc, a.s2 + a.n2);
  const double dt = *a.dt;
  double kflow[NE];
  if (active) face_at<NS, EOS, RECON, PRIMARY, LIM, 2>(g, sv, i, j, kbeg, kflow);
#pragma unroll 1
  for (int k = kbeg; k < kend; k++) {
    // the tile's i-faces and j-faces, each once, by flattened index (one face in flight per thread)
#pragma unroll 1
    for (int f = tid; f < NFI; f += WT) {
      const int r = f / (WI + 1), fi = f - r * (WI + 1);
      if (i0 + fi <= iend && j0 + r < jend) {
        double F[NE];
        face_at<NS, EOS, RECON, PRIMARY, LIM, 0>(g, sv, i0 + fi, j0 + r, k, F);
#pragma unroll
        for (int l = 0; l < NE; l++) Fi[l][f] = F[l];
      }
    }
#pragma unroll 1
    for (int f = tid; f < NFJ; f += WT) {
      const int r = f / WI, fi = f - r * WI;
      if (i0 + fi < iend && j0 + r <= jend) {
        double F[NE];
        face_at<NS, EOS, RECON, PRIMARY, LIM, 1>(g, sv, i0 + fi, j0 + r, k, F);
#pragma unroll
        for (int l = 0; l < NE; l++) Fj[l][f] = F[l];
      }
    }
    double kfhigh[NE];
    if (active) face_at<NS, EOS, RECON, PRIMARY, LIM, 2>(g, sv, i, j, k + 1, kfhigh);
    __syncthreads();
    if (active) {
      const int64_t o = Qin.off(i, j, k);
      double dQ[NE];
      {
#pragma clang fp contract(off)
        const double J = det.at(det.off(i, j, k), 0);
        const int il = row * (WI + 1) + lane, jl = row * WI + lane;
#pragma unroll
        for (int l = 0; l < NE; l++)
          dQ[l] = ((((Fi[l][il] + Fj[l][jl]) + kflow[l]) - Fi[l][il + 1]) - Fj[l][jl + WI]) - kfhigh[l];
#pragma unroll
        for (int l = 0; l < NE; l++) dQ[l] = dQ[l] * J;
      }
      if constexpr (CHEM) {
        // chemint/kernels/source.mako:3-10: the source (pga_rates) begins dQ, the divergence is added
        const int64_t orr = rates.off(i, j, k);
#pragma unroll
        for (int n = 0; n < NS - 1; n++) {
#pragma clang fp contract(off)
          dQ[5 + n] = rates.at(orr, n) + dQ[5 + n];
        }
      }
      double Qn[NE];
      {
#pragma clang fp contract(off)
        // integrators/kernels/combine.mako:1-13, terms in the reference's order
        // the terms' values are selected as values: a select of pointers would put dQ in memory
        const int64_t o0 = Q0.off(i, j, k);
#pragma unroll
        for (int l = 0; l < NE; l++) {
          const double qin = Qin.at(o, l), q0 = Q0.at(o0, l);
          double acc = 0.0;
#pragma unroll
          for (int t = 0; t < 4; t++) {
            if (t < comb.n) {
              const int src = comb.src[t];
              const double x = src == SRC_QIN ? qin : src == SRC_Q0 ? q0 : dQ[l];
              const double term = comb.withdt[t] ? comb.w[t] * dt * x : comb.w[t] * x;
              acc = t == 0 ? term : acc + term;
            }
          }
          Qn[l] = acc;
        }
      }
      EosState s;
      fromQ_regs<NS, EOS>(g, Qn, stin.at(o, 1), s);
      const int64_t oo = Qout.off(i, j, k);
#pragma unroll
      for (int l = 0; l < NE; l++) Qout.at(oo, l) = Qn[l];
      store_state(stout, stout.off(i, j, k), s);
    }
#pragma unroll
    for (int l = 0; l < NE; l++) kflow[l] = kfhigh[l];
    __syncthreads();
  }
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
  int d0, d1, d2, s0, s1, s2, e0, e1, e2;
  int zfirst;  // the first blockIdx.z of this task
};

struct HaloArgs {
  const double *K;
  View stdst, stguess;
  int ntask;
  HaloTask t[6];
};

template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_halo(HaloArgs a) {
  int n = 0;
  while (n + 1 < a.ntask && (int)blockIdx.z >= a.t[n + 1].zfirst) n++;
  const HaloTask &t = a.t[n];
  const int ii = blockIdx.x * TI + threadIdx.x, jj = blockIdx.y * TJ + threadIdx.y, kk = blockIdx.z - t.zfirst;
  if (ii >= t.e0 || jj >= t.e1 || kk >= t.e2) return;
  const int64_t od = t.Qdst.off(t.d0 + ii, t.d1 + jj, t.d2 + kk), os = t.Qsrc.off(t.s0 + ii, t.s1 + jj, t.s2 + kk);
  double Q[NS + 4];
#pragma unroll
  for (int l = 0; l < NS + 4; l++) {
    Q[l] = t.Qsrc.at(os, l);
    t.Qdst.at(od, l) = Q[l];
  }
  const Gas<NS> g{a.K};
  const int64_t ost = a.stdst.off(t.d0 + ii, t.d1 + jj, t.d2 + kk);
  EosState s;
  fromQ_regs<NS, EOS>(g, Q, a.stguess.at(ost, 1), s);
  store_state(a.stdst, ost, s);
}

// ---------------------------------------------------------------------------
// the compiled stage variants: (species, eos, recon, flux function, limiter, mechanism)
// ---------------------------------------------------------------------------
// (species, eos, recon, flux function, limiter, with a chemistry source)
#define PGA_STAGE_VARIANTS(X)                                           \
  X(9, EOS_TPG, RECON_MUSCL, FLUX_HLLC, LIM_KOREN, true)                \
  X(9, EOS_TPG, RECON_MUSCL, FLUX_HLLC, LIM_KOREN, false)               \
  X(9, EOS_TPG, RECON_MUSCL, FLUX_RUSANOV, LIM_KOREN, true)             \
  X(1, EOS_TPG, RECON_CONSTANT, FLUX_KEPAEC, LIM_NONE, false)           \
  X(1, EOS_CPG, RECON_CONSTANT, FLUX_KEPAEC, LIM_NONE, false)           \
  X(1, EOS_TPG, RECON_MUSCL, FLUX_HLLC, LIM_KOREN, false)               \
  X(2, EOS_CPG, RECON_MUSCL, FLUX_RUSANOV, LIM_KOREN, false)            \
  X(2, EOS_TPG, RECON_MUSCL, FLUX_RUSANOV, LIM_KOREN, false)

static bool matches(const pga_case_desc &d, int ns, int eos, int recon, int prim, int lim, bool chem) {
  return d.ns == ns && d.eos == eos && d.recon == recon && d.primary == prim && (recon == RECON_CONSTANT || d.limiter == lim) &&
         (d.mechanism >= 0) == chem;
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_stage_supported(const pga_case *c) {
  if (!c) return 0;
#define X(NS, EOS, RECON, PRIM, LIM, CHEM) \
  if (matches(c->desc, NS, EOS, RECON, PRIM, LIM, CHEM)) return 1;
  PGA_STAGE_VARIANTS(X)
#undef X
  return 0;
}

extern "C" int32_t pga_stage(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view Qout,
                             pga_view stout, pga_view iS, pga_view jS, pga_view kS, pga_view detJinv, pga_view rates,
                             int32_t nterms,
                             const int32_t *src, const double *w, const int32_t *withdt, const double *dt, pga_box box,
                             int32_t kchunk, void *stream) {
  if (!c || !src || !w || !withdt || !dt || nterms < 1 || nterms > 4 || kchunk < 1 || !box_ok(box)) {
    set_error("pga_stage: null argument, 1 to 4 combine terms, a k-chunk of at least 1");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int ne = d.ns + 4;
  {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    const int reach = d.recon == RECON_MUSCL ? 2 : 1;
    for (int a = 0; a < 3; a++) lo[a] = reach, hi[a] = reach;
    PGA_REQUIRE(covers(Qin, box, lo, hi, ne) && covers(stin, box, lo, hi, 8),
                "pga_stage: the stencil reaches outside Qin or its state (the halo must be %d deep)", reach);
    PGA_REQUIRE(covers(Q0, box, ne) && covers(Qout, box, ne) && covers(stout, box, 8) && covers(detJinv, box, 1),
                "pga_stage: the box reaches outside Q0, Qout, the output state or detJinv");
    const int z[3] = {0, 0, 0}, fh[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    PGA_REQUIRE(covers(iS, box, z, fh[0], 3) && covers(jS, box, z, fh[1], 3) && covers(kS, box, z, fh[2], 3),
                "pga_stage: the faces reach outside the face-area arrays");
    PGA_REQUIRE(Qin.data != Qout.data && stin.data != stout.data,
                "pga_stage: the output may not be the input (neighbouring tiles read it)");
    PGA_REQUIRE(Qin.sj == Q0.sj && Qin.sk == Q0.sk && Qin.sj == stin.sj && Qin.sk == stin.sk,
                "pga_stage: Q and state arrays must share their cell layout");
    for (int t = 0; t < nterms; t++)
      PGA_REQUIRE(src[t] >= 0 && src[t] <= 2, "pga_stage: a combine term reads Qin (0), Q0 (1) or dQ (2)");
    if (d.mechanism >= 0)
      PGA_REQUIRE(covers(rates, box, d.ns - 1), "pga_stage: a reacting case needs the rates of every cell (pga_rates)");
  }
  StageArgs a{};
  a.K = c->K;
  a.rates = view(rates);
  a.Qin = view(Qin), a.stin = view(stin), a.Q0 = view(Q0), a.Qout = view(Qout), a.stout = view(stout);
  a.S[0] = view(iS), a.S[1] = view(jS), a.S[2] = view(kS), a.det = view(detJinv);
  a.comb.n = nterms;
  for (int t = 0; t < nterms; t++) a.comb.src[t] = src[t], a.comb.w[t] = w[t], a.comb.withdt[t] = withdt[t];
  a.dt = dt;
  a.s0 = box.start[0], a.s1 = box.start[1], a.s2 = box.start[2];
  a.n0 = box.extent[0], a.n1 = box.extent[1], a.n2 = box.extent[2];
  a.kc = kchunk;
  const dim3 grid((a.n0 + WI - 1) / WI, (a.n1 + WJ - 1) / WJ, (a.n2 + kchunk - 1) / kchunk);
#define X(NS, EOS, RECON, PRIM, LIM, CHEM)                                                                  \
  if (matches(d, NS, EOS, RECON, PRIM, LIM, CHEM)) {                                                         \
    hipLaunchKernelGGL((k_stage<NS, EOS, RECON, PRIM, LIM, CHEM>), grid, dim3(WT), 0, stream_of(stream), a); \
    return launched("pga_stage");                                                                            \
  }
  PGA_STAGE_VARIANTS(X)
#undef X
  set_error("pga_stage: no fused stage compiled for this case (species %d, eos %d, recon %d, flux %d, limiter %d, "
            "mechanism %d); use the per-kernel entry points",
            d.ns, d.eos, d.recon, d.primary, d.limiter, d.mechanism);
  return PGA_EUNSUP;
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
  PGA_REQUIRE(covers(Q, box, c->desc.ns + 4) && covers(state, box, 8) && covers(rates, box, c->desc.ns - 1),
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
                            const int32_t *dst_start, const int32_t *src_start, const int32_t *extent, pga_view stdst,
                            pga_view stguess, void *stream) {
  if (!c || ntask < 0 || ntask > 6 || (ntask && (!Qdst || !Qsrc || !dst_start || !src_start || !extent))) {
    set_error("pga_halo: null argument or more than 6 slabs");
    return PGA_EINVAL;
  }
  const int ne = c->desc.ns + 4;
  HaloArgs a{};
  a.K = c->K;
  a.stdst = view(stdst);
  a.stguess = view(stguess);
  int z = 0, emax0 = 0, emax1 = 0;
  for (int n = 0; n < ntask; n++) {
    const int *ds = dst_start + 3 * n, *ss = src_start + 3 * n, *e = extent + 3 * n;
    const pga_box bd{{ds[0], ds[1], ds[2]}, {e[0], e[1], e[2]}}, bs{{ss[0], ss[1], ss[2]}, {e[0], e[1], e[2]}};
    PGA_REQUIRE(box_ok(bd) && box_ok(bs) && covers(Qdst[n], bd, ne) && covers(Qsrc[n], bs, ne) &&
                    covers(stdst, bd, 8) && covers(stguess, bd, 8),
                "pga_halo: slab %d reaches outside its arrays", n);
    PGA_REQUIRE(Qdst[n].sj == stdst.sj && Qdst[n].sk == stdst.sk, "pga_halo: Q and state must share their layout");
    HaloTask &t = a.t[n];
    t.Qdst = view(Qdst[n]), t.Qsrc = view(Qsrc[n]);
    t.d0 = ds[0], t.d1 = ds[1], t.d2 = ds[2], t.s0 = ss[0], t.s1 = ss[1], t.s2 = ss[2];
    t.e0 = e[0], t.e1 = e[1], t.e2 = e[2];
    t.zfirst = z;
    z += e[2];
    emax0 = max(emax0, e[0]);
    emax1 = max(emax1, e[1]);
  }
  a.ntask = ntask;
  if (!z || !emax0 || !emax1) return PGA_OK;
  PGA_REQUIRE(z <= 65535, "pga_halo: too many k-planes in the slabs");
  const dim3 grid((emax0 + TI - 1) / TI, (emax1 + TJ - 1) / TJ, z);
  const pga_case_desc &d = c->desc;
  bool done = false;
#define X(N)                                                                                                    \
  if (d.ns == N) {                                                                                              \
    if (d.eos == EOS_CPG)                                                                                       \
      hipLaunchKernelGGL((k_halo<N, EOS_CPG>), grid, dim3(TI, TJ), 0, stream_of(stream), a);                    \
    else                                                                                                        \
      hipLaunchKernelGGL((k_halo<N, EOS_TPG>), grid, dim3(TI, TJ), 0, stream_of(stream), a);                    \
    done = true;                                                                                                \
  }
  PGA_NS_CASES(X)
#undef X
  if (!done) {
    set_error("pga_halo: no kernel for %d species", d.ns);
    return PGA_EUNSUP;
  }
  return launched("pga_halo");
}
