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
#include "pga_diff.hpp"
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
    const unsigned t = xcd_block(), nbi = (b.e0 + 63) / 64, rest = t / nbi, nkt = (b.e2 + KTILE - 1) / KTILE;
    i = b.s0 + (int)(t % nbi) * 64 + (threadIdx.x & 63);
    k = b.s2 + (int)(rest % nkt) * KTILE + (threadIdx.x >> 6);
    j = b.s1 + (int)(rest / nkt);
    if (i >= b.s0 + b.e0 || k >= b.s2 + b.e2) return;
  } else if (!flat_cell_of(b, i, j, k)) {
    return;
  }
  flux_face<NS, EOS, RECON, PRIMARY, LIM>(K, axis, Q, st, S, F, i, j, k);
}

// every face of a block in one launch: the i faces' workgroups, then the j faces', then the k faces'
// tiles, each axis's faces taken as its own sweep takes them; each face's advective flux and, for a
// viscous case, its viscous flux added in registers (the sweeps' arithmetic to the bit), stored once
struct FacesArgs {
  const double *K;
  View Q, st, S0, S1, S2, F0, F1, F2;
  Box3 b0, b1, b2;
  unsigned n0, n1;
  double alpha;
  View tr, hsp, gr, xcc, xf0, xf1, xf2;
};

template <int NS, int EOS, int RECON, int PRIMARY, int LIM, bool VISC>
__global__ __launch_bounds__(256) void k_faces(FacesArgs a) {
  const unsigned g = xcd_block();
  int axis, i, j, k;
  if (g < a.n0) {
    axis = 0;
    if (!flat_at(a.b0, g, i, j, k)) return;
  } else if (g < a.n0 + a.n1) {
    axis = 1;
    if (!flat_at(a.b1, g - a.n0, i, j, k)) return;
  } else {
    axis = 2;
    // a column's k tiles one after another: the next tile's workgroup finds three of its seven
    // planes of cells in L2
    const Box3 &b = a.b2;
    const unsigned t = g - a.n0 - a.n1, nbi = (b.e0 + 63) / 64, rest = t / nbi, nkt = (b.e2 + KTILE - 1) / KTILE;
    i = b.s0 + (int)(t % nbi) * 64 + (threadIdx.x & 63);
    k = b.s2 + (int)(rest % nkt) * KTILE + (threadIdx.x >> 6);
    j = b.s1 + (int)(rest / nkt);
    if (i >= b.s0 + b.e0 || k >= b.s2 + b.e2) return;
  }
  const View &S = axis == 0 ? a.S0 : axis == 1 ? a.S1 : a.S2;
  const View &F = axis == 0 ? a.F0 : axis == 1 ? a.F1 : a.F2;
  double Fv[NS + 4];
  flux_at<NS, EOS, RECON, PRIMARY, LIM>(a.K, axis, a.Q, a.st, S, i, j, k, Fv);
  if constexpr (VISC) {
    const View &xf = axis == 0 ? a.xf0 : axis == 1 ? a.xf1 : a.xf2;
    diff_face<NS>(a.Q, a.st, a.tr, a.hsp, a.gr, a.xcc, xf, S, a.alpha, axis, i, j, k, Fv);
  }
  // the face array streams past the caches: the stencil's Q and state keep L2 (5% on the 128^3 step's
  // faces); the stage reads the array back from HBM either way
  const int64_t oF = F.off(i, j, k);
#pragma unroll
  for (int l = 0; l < NS + 4; l++) __builtin_nontemporal_store(Fv[l], &F.at(oF, l));
}

template <int NS, int EOS>
static void launch_faces(const pga_case *c, const FacesArgs &a, bool visc, dim3 grid, hipStream_t s) {
  const int lim = c->desc.limiter, recon = c->desc.recon, prim = c->desc.primary;
#define L(RECON, PRIM, LIM)                                                                                    \
  if (recon == RECON && prim == PRIM && (RECON == RECON_CONSTANT || lim == LIM)) {                            \
    if (visc)                                                                                                  \
      hipLaunchKernelGGL((k_faces<NS, EOS, RECON, PRIM, LIM, true>), grid, dim3(FLAT), 0, s, a);               \
    else                                                                                                       \
      hipLaunchKernelGGL((k_faces<NS, EOS, RECON, PRIM, LIM, false>), grid, dim3(FLAT), 0, s, a);              \
  }
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

extern "C" int32_t pga_faces(const pga_case *c, pga_view Q, pga_view state, const pga_view *S, const pga_view *F,
                             const pga_box *boxes, int32_t viscous, double alpha, pga_view trstate, pga_view hsp,
                             pga_view grads, pga_view xcc, const pga_view *xf, void *stream) {
  if (!c || !S || !F || !boxes || (viscous && !xf)) {
    set_error("pga_faces: null case, face areas, fluxes, boxes, or (viscous) face centres");
    return PGA_EINVAL;
  }
  const pga_case_desc &d = c->desc;
  if (d.recon < 0 || d.recon > 1 || d.primary < 0 || d.primary > 2 || (d.recon == 1 && (d.limiter < 1 || d.limiter > 5))) {
    set_error("pga_faces: no reconstruction %d / flux function %d / limiter %d", d.recon, d.primary, d.limiter);
    return PGA_EUNSUP;
  }
  const int ne = d.ns + 4;
  for (int axis = 0; axis < 3; axis++) {
    const pga_box &box = boxes[axis];
    PGA_REQUIRE(box_ok(box), "pga_faces: the faces of axis %d have a negative extent", axis);
    if (empty(box)) continue;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    lo[axis] = d.recon == 1 ? 2 : 1;
    hi[axis] = d.recon == 1 ? 1 : 0;
    PGA_REQUIRE(covers(Q, box, lo, hi, ne) && covers(state, box, lo, hi, nstate(d)),
                "pga_faces: the stencil of axis %d reaches outside Q or state", axis);
    PGA_REQUIRE(covers(S[axis], box, 3) && covers(F[axis], box, ne), "pga_faces: the faces of axis %d reach outside S or F",
                axis);
    if (viscous) {
      int vl[3] = {0, 0, 0};
      const int vh[3] = {0, 0, 0};
      vl[axis] = 1;
      PGA_REQUIRE(covers(trstate, box, vl, vh, 2 + d.ns) && covers(hsp, box, vl, vh, d.ns) &&
                      covers(grads, box, vl, vh, 3 * (ne - 1)) && covers(xcc, box, vl, vh, 3) && covers(xf[axis], box, 3),
                  "pga_faces: the viscous face of axis %d reaches outside a cell array or xf", axis);
      PGA_REQUIRE(Q.sj == state.sj && Q.sk == state.sk && Q.sj == trstate.sj && Q.sk == trstate.sk && Q.sj == hsp.sj &&
                      Q.sk == hsp.sk && Q.sj == grads.sj && Q.sk == grads.sk && Q.sj == xcc.sj && Q.sk == xcc.sk,
                  "pga_faces: the cell arrays must share their layout");
    }
    if (!flat_fits(box)) return PGA_ERANGE;
  }
  FacesArgs a{};
  a.K = c->K;
  a.Q = view(Q), a.st = view(state);
  a.S0 = view(S[0]), a.S1 = view(S[1]), a.S2 = view(S[2]);
  a.F0 = view(F[0]), a.F1 = view(F[1]), a.F2 = view(F[2]);
  a.b0 = box3(boxes[0]), a.b1 = box3(boxes[1]), a.b2 = box3(boxes[2]);
  a.n0 = empty(boxes[0]) ? 0 : flat_grid(a.b0).x;
  a.n1 = empty(boxes[1]) ? 0 : flat_grid(a.b1).x;
  const unsigned n2 = empty(boxes[2]) ? 0 : ktile_grid(a.b2).x;
  if (viscous) {
    a.alpha = alpha;
    a.tr = view(trstate), a.hsp = view(hsp), a.gr = view(grads), a.xcc = view(xcc);
    a.xf0 = view(xf[0]), a.xf1 = view(xf[1]), a.xf2 = view(xf[2]);
  }
  const unsigned total = a.n0 + a.n1 + n2;
  if (!total) return PGA_OK;
  const bool done = dispatch_eos(d, [&]<int N, int E>() { launch_faces<N, E>(c, a, viscous != 0, dim3(total), stream_of(stream)); });
  if (!done) {
    set_error("pga_faces: no kernel for %d species with eos %d", d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_faces");
}
