// The k sweep and the fused stage in one launch (pga_stage_march): a thread walks the cells of its
// (i, j) column over a chunk of k-planes, computes each k face's flux once and keeps it in the
// workgroup's LDS for the two cells it lies between, and runs the fused stage's divergence, source and
// RK combine on each cell; the EOS of the result follows as in pga_stage. The k-face flux array of the
// split path (written by one sweep, read back by the stage: 2 x 211 MiB a stage at 128^3) is never
// stored. Every value is the split path's to the bit: the same face function and the same order of
// sums.
#include <type_traits>

#include "pga_internal.hpp"
#include "pga_stagecore.hpp"

namespace pga {

template <int NS, int EOS, int RECON, int PRIMARY, int LIM>
__global__ __launch_bounds__(WT) void k_march(StageArgs a, View kS, int chem) {
  __shared__ double lds[2 * (NS + 4) * WT];
  const int i = a.s0 + blockIdx.x * WI + (threadIdx.x & (WI - 1));
  const int j = a.s1 + blockIdx.y * WJ + threadIdx.x / WI;
  if (i >= a.s0 + a.n0 || j >= a.s1 + a.n1) return;
  const int kbeg = a.s2 + blockIdx.z * a.kc, kend = min(kbeg + a.kc, a.s2 + a.n2);
  if (chem)
    march_cells<NS, EOS, RECON, PRIMARY, LIM, true>(a, kS, lds, i, j, kbeg, kend);
  else
    march_cells<NS, EOS, RECON, PRIMARY, LIM, false>(a, kS, lds, i, j, kbeg, kend);
}

template <int NS, int EOS>
static void launch_march(const pga_case *c, const StageArgs &a, const View &kS, int chem, dim3 grid, hipStream_t s) {
  const int lim = c->desc.limiter, recon = c->desc.recon, prim = c->desc.primary;
#define L(RECON, PRIM, LIM)                                                                         \
  if (recon == RECON && prim == PRIM && (RECON == RECON_CONSTANT || lim == LIM))                   \
    hipLaunchKernelGGL((k_march<NS, EOS, RECON, PRIM, LIM>), grid, dim3(WT), 0, s, a, kS, chem);
  L(RECON_CONSTANT, FLUX_KEPAEC, LIM_NONE)
  L(RECON_CONSTANT, FLUX_HLLC, LIM_NONE)
  L(RECON_CONSTANT, FLUX_RUSANOV, LIM_NONE)
#define M(PRIM) L(RECON_MUSCL, PRIM, LIM_KOREN) L(RECON_MUSCL, PRIM, LIM_MINMOD) L(RECON_MUSCL, PRIM, LIM_MC) \
  L(RECON_MUSCL, PRIM, LIM_SUPERBEE) L(RECON_MUSCL, PRIM, LIM_VANLEER)
  M(FLUX_HLLC)
  M(FLUX_RUSANOV)
  M(FLUX_KEPAEC)
#undef M
#undef L
}

template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_march_eos(const double *__restrict__ K, View Q, View stin, View stout, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  eos_cell<NS, EOS>(K, Q, stin, stout, i, j, k);
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_stage_march(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view Qout,
                                   pga_view stout, pga_view iF, pga_view jF, pga_view kS, pga_view detJinv,
                                   pga_view rates, int32_t nterms, const int32_t *src, const double *w,
                                   const int32_t *withdt, const double *dt, pga_box box, int32_t kchunk,
                                   void *stream) {
  const char *who = "pga_stage_march";
  if (!c || !src || !w || !withdt || !dt || nterms < 1 || nterms > 4 || kchunk < 1 || !box_ok(box)) {
    set_error("%s: null argument, 1 to 4 combine terms, a k-chunk of at least 1", who);
    return PGA_EINVAL;
  }
  const pga_case_desc &d = c->desc;
  if (d.recon < 0 || d.recon > 1 || d.primary < 0 || d.primary > 2 || (d.recon == 1 && (d.limiter < 1 || d.limiter > 5))) {
    set_error("%s: no reconstruction %d / flux function %d / limiter %d", who, d.recon, d.primary, d.limiter);
    return PGA_EUNSUP;
  }
  if (Qin.data == Qout.data) {
    set_error("%s: the output is the input; a chunk's last k face reads cells the next chunk writes (use pga_flux "
              "along k and pga_stage)", who);
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  const int ne = d.ns + 4;
  const bool chem = d.mechanism >= 0;
  {
    PGA_REQUIRE(covers(Qin, box, ne) && covers(stin, box, nstate(d)) && covers(Q0, box, ne) && covers(Qout, box, ne) &&
                    covers(stout, box, nstate(d)) && covers(detJinv, box, 1),
                "%s: the box reaches outside Qin, Q0, Qout, a state or detJinv", who);
    // the k faces' stencil along k: LL, L, R, RR under MUSCL, L and R otherwise, for faces kbeg..kend
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    lo[2] = d.recon == 1 ? 2 : 1;
    hi[2] = d.recon == 1 ? 2 : 1;
    PGA_REQUIRE(covers(Qin, box, lo, hi, ne) && covers(stin, box, lo, hi, nstate(d)),
                "%s: the k faces' stencil reaches outside Qin or its state", who);
    const int z[3] = {0, 0, 0}, fh[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    PGA_REQUIRE(covers(iF, box, z, fh[0], ne) && covers(jF, box, z, fh[1], ne) && covers(kS, box, z, fh[2], 3),
                "%s: the box's faces reach outside iF, jF or kS", who);
    for (int t = 0; t < nterms; t++)
      PGA_REQUIRE(src[t] >= 0 && src[t] <= 2, "%s: a combine term reads Qin (0), Q0 (1) or dQ (2)", who);
    if (chem)
      PGA_REQUIRE(covers(rates, box, d.ns - 1), "%s: a reacting case needs the rates of every cell (pga_rates)", who);
  }
  StageArgs a{};
  a.K = c->K;
  a.rates = view(rates);
  a.Qin = view(Qin), a.stin = view(stin), a.Q0 = view(Q0), a.Qout = view(Qout), a.stout = view(stout);
  a.iF = view(iF), a.jF = view(jF), a.det = view(detJinv);
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
  const View kSv = view(kS);
  const bool done = dispatch_eos(d, [&]<int N, int E>() {
    launch_march<N, E>(c, a, kSv, chem ? 1 : 0, grid, stream_of(stream));
    hipLaunchKernelGGL((k_march_eos<N, E>), grid_of(b), dim3(TI, TJ), 0, stream_of(stream), c->K, a.Qout, a.stin,
                       a.stout, b);
  });
  if (!done) {
    set_error("%s: no kernel for %d species with eos %d", who, d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched(who);
}
