// The advective face flux of the per-kernel path: one launch per axis over the
// faces of a box, one face per thread, lanes along the contiguous i index so the
// stencil reads of every axis are coalesced.
//
// The i and j faces take a flat launch (rows of faces back to back, XCD-ordered
// blocks): the rows a j-face stencil spans are shared with the blocks dispatched
// beside it on the same XCD, so L2 serves the reuse. A k-face stencil spans four
// planes, further apart than an XCD's blocks in flight reach: there a workgroup
// takes a 64 (i) x 4 (k) tile of faces, its four waves four consecutive k-faces,
// and the seven planes of cells they span are read once into L1 and L2 for all
// four rather than four times.
#include "pga_flux.hpp"
#include "pga_internal.hpp"

namespace pga {

constexpr int KTILE = 4;

__host__ inline dim3 ktile_grid(const Box3 &b) {
  return dim3((unsigned)(((b.e0 + 63) / 64) * b.e1 * ((b.e2 + KTILE - 1) / KTILE)));
}

template <int NS, int EOS, int RECON, int PRIMARY, int LIM>
__global__ __launch_bounds__(256) void k_flux(const double *__restrict__ K, int axis, View Q, View st, View S,
                                              View F, Box3 b) {
  int i, j, k;
  if (axis == 2) {
    const unsigned t = xcd_block(), nbi = (b.e0 + 63) / 64, rest = t / nbi;
    i = b.s0 + (int)(t % nbi) * 64 + (threadIdx.x & 63);
    j = b.s1 + (int)(rest % b.e1);
    k = b.s2 + (int)(rest / b.e1) * KTILE + (threadIdx.x >> 6);
    if (i >= b.s0 + b.e0 || k >= b.s2 + b.e2) return;
  } else if (!flat_cell_of(b, i, j, k)) {
    return;
  }
  flux_face<NS, EOS, RECON, PRIMARY, LIM>(K, axis, Q, st, S, F, i, j, k);
}

template <int NS, int EOS>
static void launch_flux(const pga_case *c, int axis, pga_view Q, pga_view st, pga_view S, pga_view F, Box3 b,
                        hipStream_t s) {
  const int lim = c->desc.limiter, recon = c->desc.recon, prim = c->desc.primary;
#define L(RECON, PRIM, LIM)                                                                                    \
  if (recon == RECON && prim == PRIM && (RECON == RECON_CONSTANT || lim == LIM))                              \
    flat_slabs(b, [&](const Box3 &sb) {                                                                        \
      hipLaunchKernelGGL((k_flux<NS, EOS, RECON, PRIM, LIM>), axis == 2 ? ktile_grid(sb) : flat_grid(sb),       \
                         dim3(FLAT), 0, s, c->K, axis, view(Q), view(st), view(S), view(F), sb);                \
    });
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

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_flux(const pga_case *c, int32_t axis, pga_view Q, pga_view state, pga_view S, pga_view F,
                            pga_box box, void *stream) {
  if (!c || axis < 0 || axis > 2 || !box_ok(box)) {
    set_error("pga_flux: null case, bad axis or box");
    return PGA_EINVAL;
  }
  const pga_case_desc &d = c->desc;
  if (d.recon < 0 || d.recon > 1 || d.primary < 0 || d.primary > 2 || (d.recon == 1 && (d.limiter < 1 || d.limiter > 5))) {
    set_error("pga_flux: no reconstruction %d / flux function %d / limiter %d", d.recon, d.primary, d.limiter);
    return PGA_EUNSUP;
  }
  if (empty(box)) return PGA_OK;
  {
    // a face's stencil: LL, L, R, RR along the axis under MUSCL, L and R otherwise
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    lo[axis] = d.recon == 1 ? 2 : 1;
    hi[axis] = d.recon == 1 ? 1 : 0;
    PGA_REQUIRE(covers(Q, box, lo, hi, d.ns + 4) && covers(state, box, lo, hi, nstate(d)),
                "pga_flux: the faces' stencil reaches outside Q or state");
    PGA_REQUIRE(covers(S, box, 3) && covers(F, box, d.ns + 4), "pga_flux: the faces reach outside S or F");
  }
  if (!flat_fits(box)) return PGA_ERANGE;
  const Box3 b = box3(box);
  const bool done = dispatch_eos(d, [&]<int N, int E>() { launch_flux<N, E>(c, axis, Q, state, S, F, b, stream_of(stream)); });
  if (!done) {
    set_error("pga_flux: no kernel for %d species with eos %d", d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_flux");
}
