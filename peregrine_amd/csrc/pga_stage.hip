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

#include "pga_internal.hpp"
#include "pga_stagecore.hpp"

namespace pga {

// the divergence, the source and the RK combine of a column of cells: memory-bound
template <int NE, bool CHEM>
__global__ __launch_bounds__(WT) void k_stage(StageArgs a) {
  const int i = a.s0 + blockIdx.x * WI + (threadIdx.x & (WI - 1));
  const int j = a.s1 + blockIdx.y * WJ + threadIdx.x / WI;
  if (i >= a.s0 + a.n0 || j >= a.s1 + a.n1) return;
  const int kbeg = a.s2 + blockIdx.z * a.kc;
  stage_cells<NE, CHEM>(a, i, j, kbeg, min(kbeg + a.kc, a.s2 + a.n2));
}

// a small block's stage and the EOS of its result in one launch: each thread's cells after its
// column's combine (a block of a few thousand cells is bound by launches, not by the registers the
// two phases hold at once, which at 128^3 favour the two launches apart)
template <int NS, int EOS, bool CHEM>
__global__ __launch_bounds__(WT) void k_stage_whole(StageArgs a) {
  const int i = a.s0 + blockIdx.x * WI + (threadIdx.x & (WI - 1));
  const int j = a.s1 + blockIdx.y * WJ + threadIdx.x / WI;
  if (i >= a.s0 + a.n0 || j >= a.s1 + a.n1) return;
  const int kbeg = a.s2 + blockIdx.z * a.kc, kend = min(kbeg + a.kc, a.s2 + a.n2);
  if constexpr (NS + 4 <= 8) {
    // a few rows: each cell's reads at once, its EOS from the new Q in registers
    const Gas<NS> g{a.K};
    const double dt = *a.dt;
    for (int k = kbeg; k < kend; k++) {
      const double Tguess = a.stin.at(a.stin.off(i, j, k), 1);
      double Qn[NS + 4];
      stage_cell_regs<NS + 4, CHEM>(a, i, j, k, dt, Qn);
      EosState s;
      fromQ_regs<NS, EOS>(g, Qn, Tguess, s);
      store_state<EOS>(a.stout, a.stout.off(i, j, k), s);
    }
  } else {
    stage_cells<NS + 4, CHEM>(a, i, j, kbeg, kend);
    for (int k = kbeg; k < kend; k++) eos_cell<NS, EOS>(a.K, a.Qout, a.stin, a.stout, i, j, k);
  }
}

// the EOS of the stage's result, T started from the input state (stin may be stout)
template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_stage_eos(const double *__restrict__ K, View Q, View stin, View stout, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  eos_cell<NS, EOS>(K, Q, stin, stout, i, j, k);
}

// the carried species' production rates of each cell (chemint/kernels/cell.mako:5-20), the
// source a reacting stage begins dQ with
template <class M>
__global__ __launch_bounds__(256) void k_rates(const double *__restrict__ K, const double *__restrict__ C, View Q,
                                               View st, View rates, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  rates_cell<M>(K, C, Q, st, rates, i, j, k);
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

// one thread per halo cell over the slabs back to back, i fastest within a slab; each slab begins on
// a workgroup's boundary, so a workgroup's slab is a scalar choice
template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_halo(HaloArgs a) {
  const int64_t c0 = (int64_t)blockIdx.x * 256;
  int n = 0;
  while (n + 1 < a.ntask && c0 >= a.t[n + 1].first) n++;
  const HaloTask &t = a.t[n];
  // a slab holds fewer than 2^31 cells
  const unsigned r = (unsigned)(c0 - t.first) + threadIdx.x;
  if (r >= t.cells) return;
  const unsigned e0 = t.e0, e1 = t.e1;
  const int ii = (int)(r % e0), jj = (int)((r / e0) % e1), kk = (int)(r / (e0 * e1));
  const int64_t os = t.sbase + ii * t.st0 + jj * t.st1 + kk * t.st2;
  halo_cell<NS, EOS>(a.K, t.Qdst, t.Qsrc, os, t.turn, a.stdst, a.stguess, t.d0 + ii, t.d1 + jj, t.d2 + kk);
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

// a stage with one combine (nA = 0) or two, the first into QA
static int32_t stage_impl(const char *who, const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view QA,
                          pga_view Qout, pga_view stout, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv,
                          pga_view rates, int32_t nA, const int32_t *srcA, const double *wA, const int32_t *withdtA,
                          int32_t nterms, const int32_t *src, const double *w, const int32_t *withdt, const double *dt,
                          pga_box box, int32_t kchunk, void *stream) {
  if (!c || !src || !w || !withdt || !dt || nterms < 1 || nterms > 4 || kchunk < 1 || !box_ok(box) || nA < 0 ||
      nA > 4 || (nA && (!srcA || !wA || !withdtA))) {
    set_error("%s: null argument, 1 to 4 combine terms (0 to 4 in the first), a k-chunk of at least 1", who);
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int ne = d.ns + 4;
  const bool chem = d.mechanism >= 0;
  {
    PGA_REQUIRE(covers(Qin, box, ne) && covers(stin, box, nstate(d)) && covers(Q0, box, ne) && covers(Qout, box, ne) &&
                    covers(stout, box, nstate(d)) && covers(detJinv, box, 1),
                "%s: the box reaches outside Qin, Q0, Qout, a state or detJinv", who);
    if (nA) PGA_REQUIRE(covers(QA, box, ne), "%s: the box reaches outside the first combine's register", who);
    const int z[3] = {0, 0, 0}, fh[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    PGA_REQUIRE(covers(iF, box, z, fh[0], ne) && covers(jF, box, z, fh[1], ne) && covers(kF, box, z, fh[2], ne),
                "%s: the box's faces reach outside the face flux arrays", who);
    for (int t = 0; t < nterms; t++)
      PGA_REQUIRE(src[t] >= 0 && src[t] <= (nA ? 4 : 2) && src[t] != SRC_OWNA,
                  "%s: a combine term reads Qin (0), Q0 (1), dQ (2) or the first combine's value (4)", who);
    for (int t = 0; t < nA; t++)
      PGA_REQUIRE(srcA[t] >= 0 && srcA[t] <= 3,
                  "%s: a first-combine term reads Qin (0), Q0 (1), dQ (2) or its own register (3)", who);
    if (chem)
      PGA_REQUIRE(covers(rates, box, d.ns - 1), "%s: a reacting case needs the rates of every cell (pga_rates)", who);
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
  if (nA) {
    a.two = 1;
    a.QA = view(QA);
    a.combA.n = nA;
    for (int t = 0; t < nA; t++) {
      a.combA.src[t] = srcA[t], a.combA.w[t] = wA[t], a.combA.withdt[t] = withdtA[t];
      a.q0 |= srcA[t] == SRC_Q0;
    }
  }
  a.dt = dt;
  a.s0 = box.start[0], a.s1 = box.start[1], a.s2 = box.start[2];
  a.n0 = box.extent[0], a.n1 = box.extent[1], a.n2 = box.extent[2];
  a.kc = kchunk;
  const dim3 grid((a.n0 + WI - 1) / WI, (a.n1 + WJ - 1) / WJ, (a.n2 + kchunk - 1) / kchunk);
  const Box3 b = box3(box);
  // a block of at most 2^16 cells (up to 40^3) in one launch, a larger one in two
  const bool whole = (int64_t)a.n0 * a.n1 * a.n2 <= (int64_t(1) << 16);
  const bool done = dispatch_eos(d, [&]<int N, int E>() {
    if (whole) {
      if (chem)
        hipLaunchKernelGGL((k_stage_whole<N, E, true>), grid, dim3(WT), 0, stream_of(stream), a);
      else
        hipLaunchKernelGGL((k_stage_whole<N, E, false>), grid, dim3(WT), 0, stream_of(stream), a);
      return;
    }
    if (chem)
      hipLaunchKernelGGL((k_stage<N + 4, true>), grid, dim3(WT), 0, stream_of(stream), a);
    else
      hipLaunchKernelGGL((k_stage<N + 4, false>), grid, dim3(WT), 0, stream_of(stream), a);
    hipLaunchKernelGGL((k_stage_eos<N, E>), grid_of(b), dim3(TI, TJ), 0, stream_of(stream), c->K, a.Qout, a.stin,
                       a.stout, b);
  });
  if (!done) {
    set_error("%s: no kernel for %d species with eos %d", who, d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched(who);
}

extern "C" int32_t pga_stage(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view Qout,
                             pga_view stout, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv, pga_view rates,
                             int32_t nterms, const int32_t *src, const double *w, const int32_t *withdt,
                             const double *dt, pga_box box, int32_t kchunk, void *stream) {
  return stage_impl("pga_stage", c, Qin, stin, Q0, pga_view{}, Qout, stout, iF, jF, kF, detJinv, rates, 0, nullptr,
                    nullptr, nullptr, nterms, src, w, withdt, dt, box, kchunk, stream);
}

extern "C" int32_t pga_stage2(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view QA,
                              pga_view Qout, pga_view stout, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv,
                              pga_view rates, int32_t nA, const int32_t *srcA, const double *wA,
                              const int32_t *withdtA, int32_t nterms, const int32_t *src, const double *w,
                              const int32_t *withdt, const double *dt, pga_box box, int32_t kchunk, void *stream) {
  if (nA < 1) {
    set_error("pga_stage2: the first combine needs 1 to 4 terms");
    return PGA_EINVAL;
  }
  return stage_impl("pga_stage2", c, Qin, stin, Q0, QA, Qout, stout, iF, jF, kF, detJinv, rates, nA, srcA, wA,
                    withdtA, nterms, src, w, withdt, dt, box, kchunk, stream);
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
    PGA_REQUIRE(t.cells < FLAT_MAX, "pga_halo: slab %d has %lld cells; at most 2^31", n, (long long)t.cells);
    total += (t.cells + 255) / 256 * 256;
  }
  a.ntask = ntask;
  if (!total) return PGA_OK;
  const int64_t nblk = total / 256;
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
