// Cell kernels of the per-kernel path: the EOS inversion, initialisation, the flux
// divergence, the RK combine, the periodic halo fill and the chemistry source.
//
// Launch shape: a 64 x 4 block, the 64 lanes of a wave along the contiguous i
// index so every component plane is read and written in whole 512-byte runs;
// blockIdx.z steps k.
#include "generated/mechanisms.hpp"
#include "pga_chem.hpp"
#include "pga_eos.hpp"
#include "pga_internal.hpp"

namespace pga {

template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_fromQ(const double *__restrict__ K, View Q, View st, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  cell_fromQ<NS, EOS>(Gas<NS>{K}, Q, st, Q.off(i, j, k));
}

// euler/kernels/state/fromprims.mako:4-13 with solution.mako:2-23
template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_fromprims(const double *__restrict__ K, View P, View Q, View st, Box3 b) {
#pragma clang fp contract(off)
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  const Gas<NS> g{K};
  const int64_t o = Q.off(i, j, k);
  double Y[NS > 1 ? NS - 1 : 1];
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Y[n] = P.at(o, 5 + n);
  const double u0 = P.at(o, 1), u1 = P.at(o, 2), u2 = P.at(o, 3);
  MixMom<NS, EOS> m;
  mixmom<NS, EOS>(g, [&](int n) { return Y[n]; }, m);
  EosState s;
  frompT<NS, EOS>(g, m, [&](int n) { return Y[n]; }, P.at(o, 0), P.at(o, 4), s);
  Q.at(o, 0) = s.rho;
  Q.at(o, 1) = s.rho * u0;
  Q.at(o, 2) = s.rho * u1;
  Q.at(o, 3) = s.rho * u2;
  Q.at(o, 4) = s.rho * (s.e + 0.5 * (u0 * u0 + u1 * u1 + u2 * u2));
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Q.at(o, 5 + n) = s.rho * Y[n];
  store_state<EOS>(st, o, s);
}

// advec/kernels/applyflux.mako:2-11, the sum in the reference's order
__global__ __launch_bounds__(256) void k_applyflux(int ne, View dQ, View iF, View jF, View kF, View J, int append,
                                                   Box3 b) {
#pragma clang fp contract(off)
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  const int64_t oi = iF.off(i, j, k), oj = jF.off(i, j, k), ok = kF.off(i, j, k), oq = dQ.off(i, j, k);
  const double det = J.d[J.off(i, j, k)];
  for (int l = 0; l < ne; l++) {
    const double v = (iF.at(oi, l) + jF.at(oj, l) + kF.at(ok, l) - iF.at(oi + 1, l) - jF.at(oj + jF.sj, l) -
                      kF.at(ok + kF.sk, l)) * det;
    if (append)
      dQ.at(oq, l) += v;
    else
      dQ.at(oq, l) = v;
  }
}

struct CombineArgs {
  View X[4];
  double w[4];
};

// integrators/kernels/combine.mako:1-13
__global__ __launch_bounds__(256) void k_combine(int ne, View out, int own, double wout, int nterms, CombineArgs a,
                                                 const double *__restrict__ dt, int step, Box3 b) {
#pragma clang fp contract(off)
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  const double h = step ? *dt : 1.0;
  const int64_t o = out.off(i, j, k);
  for (int l = 0; l < ne; l++) {
    double acc = 0.0;
    bool first = true;
    if (own) {
      acc = wout * out.at(o, l);
      first = false;
    }
    for (int m = 0; m < nterms; m++) {
      const View &X = a.X[m];
      const double x = X.at(X.off(i, j, k), l);
      const double term = (step && m == nterms - 1) ? a.w[m] * h * x : a.w[m] * x;
      acc = first ? term : acc + term;
      first = false;
    }
    out.at(o, l) = acc;
  }
}

// place.mako:11-37 on a self-periodic pair: the low halo layer l (outward) from interior plane
// ng - 1 - l + n, the high halo layer l from ng + l; lanes along i, whatever the axis
__global__ __launch_bounds__(256) void k_fill_periodic(View A, int ncomp, int n0, int n1, int n2, int ng, int depth,
                                                       int axis) {
  const int n[3] = {n0, n1, n2};
  int ext[3] = {n0, n1, n2};
  ext[axis] = depth;
  const int ii = blockIdx.x * TI + threadIdx.x, jj = blockIdx.y * TJ + threadIdx.y;
  const int kk = blockIdx.z % ext[2], comp = blockIdx.z / ext[2];
  if (ii >= ext[0] || jj >= ext[1] || comp >= ncomp) return;
  int h[3] = {ii + ng, jj + ng, kk + ng};
  const int layer = (axis == 0 ? ii : axis == 1 ? jj : kk);
  int dl[3] = {h[0], h[1], h[2]}, sl[3] = {h[0], h[1], h[2]}, dh[3] = {h[0], h[1], h[2]}, sh[3] = {h[0], h[1], h[2]};
  dl[axis] = ng - 1 - layer;
  sl[axis] = ng - 1 - layer + n[axis];
  dh[axis] = ng + n[axis] + layer;
  sh[axis] = ng + layer;
  A.at(A.off(dl[0], dl[1], dl[2]), comp) = A.at(A.off(sl[0], sl[1], sl[2]), comp);
  A.at(A.off(dh[0], dh[1], dh[2]), comp) = A.at(A.off(sh[0], sh[1], sh[2]), comp);
}

template <class M>
__global__ __launch_bounds__(256) void k_source(const double *__restrict__ K, const double *__restrict__ C, View Q,
                                                View st, View dQ, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  cell_source<M>(Gas<M::NS>{K}, C, Q, st, dQ, Q.off(i, j, k));
}

template <class M>
__global__ __launch_bounds__(256) void k_production(const double *__restrict__ K, const double *__restrict__ C, int n,
                                                    const double *__restrict__ rho, const double *__restrict__ Y,
                                                    const double *__restrict__ T, double *__restrict__ omega) {
  constexpr int NS = M::NS;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double y[NS], w[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) y[s] = Y[(int64_t)s * n + p];
  M::production(Gas<NS>{K}, C, rho[p], y, T[p], w);
#pragma unroll
  for (int s = 0; s < NS; s++) omega[(int64_t)s * n + p] = w[s];
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_fromQ(const pga_case *c, pga_view Q, pga_view state, const pga_box *boxes, int32_t nboxes,
                             void *stream) {
  if (!c || nboxes < 0 || (nboxes && !boxes)) {
    set_error("pga_fromQ: null case or boxes");
    return PGA_EINVAL;
  }
  for (int n = 0; n < nboxes; n++) {
    if (!box_ok(boxes[n])) {
      set_error("pga_fromQ: box %d has a negative start or extent, or more than 65535 k-planes", n);
      return PGA_EINVAL;
    }
    if (empty(boxes[n])) continue;
    PGA_REQUIRE(covers(Q, boxes[n], c->desc.ns + 4) && covers(state, boxes[n], nstate(c->desc)),
                "pga_fromQ: box %d reaches outside Q or state, or they hold too few components", n);
    const Box3 b = box3(boxes[n]);
    const dim3 grid = grid_of(b), block(TI, TJ);
    const bool done = dispatch_eos(c->desc, [&]<int N, int E>() {
      hipLaunchKernelGGL((k_fromQ<N, E>), grid, block, 0, stream_of(stream), c->K, view(Q), view(state), b);
    });
    if (!done) {
      set_error("pga_fromQ: no kernel for %d species with eos %d", c->desc.ns, c->desc.eos);
      return PGA_EUNSUP;
    }
    const int32_t rc = launched("pga_fromQ");
    if (rc) return rc;
  }
  return PGA_OK;
}

extern "C" int32_t pga_fromprims(const pga_case *c, pga_view prims, pga_view Q, pga_view state, pga_box box,
                                 void *stream) {
  if (!c || !box_ok(box)) {
    set_error("pga_fromprims: null case or bad box");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  PGA_REQUIRE(covers(prims, box, c->desc.ns + 4) && covers(Q, box, c->desc.ns + 4) && covers(state, box, nstate(c->desc)),
              "pga_fromprims: the box reaches outside prims, Q or state");
  const Box3 b = box3(box);
  const dim3 grid = grid_of(b), block(TI, TJ);
  const bool done = dispatch_eos(c->desc, [&]<int N, int E>() {
    hipLaunchKernelGGL((k_fromprims<N, E>), grid, block, 0, stream_of(stream), c->K, view(prims), view(Q), view(state),
                       b);
  });
  if (!done) {
    set_error("pga_fromprims: no kernel for %d species with eos %d", c->desc.ns, c->desc.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_fromprims");
}

extern "C" int32_t pga_applyflux(int32_t ne, pga_view dQ, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv,
                                 int32_t append, pga_box box, void *stream) {
  if (ne <= 0 || !box_ok(box)) {
    set_error("pga_applyflux: bad ne or box");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  {
    const int z[3] = {0, 0, 0}, hi[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    PGA_REQUIRE(covers(dQ, box, ne) && covers(detJinv, box, 1) && covers(iF, box, z, hi[0], ne) &&
                    covers(jF, box, z, hi[1], ne) && covers(kF, box, z, hi[2], ne),
                "pga_applyflux: the box (and its high faces) reaches outside dQ, detJinv or a face array");
  }
  const Box3 b = box3(box);
  hipLaunchKernelGGL(k_applyflux, grid_of(b), dim3(TI, TJ), 0, stream_of(stream), ne, view(dQ), view(iF), view(jF),
                     view(kF), view(detJinv), append, b);
  return launched("pga_applyflux");
}

extern "C" int32_t pga_combine(int32_t ne, pga_view out, int32_t own, double wout, int32_t nterms, const pga_view *X,
                               const double *w, const double *dt, int32_t step, pga_box box, void *stream) {
  if (ne <= 0 || nterms < 0 || nterms > 4 || (nterms && (!X || !w)) || (step && !dt) || !box_ok(box)) {
    set_error("pga_combine: bad arguments (at most 4 terms; a step needs dt)");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  PGA_REQUIRE(covers(out, box, ne), "pga_combine: the box reaches outside out");
  for (int m = 0; m < nterms; m++) PGA_REQUIRE(covers(X[m], box, ne), "pga_combine: the box reaches outside X[%d]", m);
  CombineArgs a{};
  for (int m = 0; m < nterms; m++) {
    a.X[m] = view(X[m]);
    a.w[m] = w[m];
  }
  const Box3 b = box3(box);
  hipLaunchKernelGGL(k_combine, grid_of(b), dim3(TI, TJ), 0, stream_of(stream), ne, view(out), own, wout, nterms, a,
                     dt, step, b);
  return launched("pga_combine");
}

extern "C" int32_t pga_fill_periodic(pga_view A, int32_t ncomp, const int32_t *ncell, int32_t ng, int32_t depth,
                                     int32_t axis, void *stream) {
  if (!ncell || axis < 0 || axis > 2 || depth < 0 || depth > ng || ncomp <= 0) {
    set_error("pga_fill_periodic: bad axis, depth or component count");
    return PGA_EINVAL;
  }
  if (depth > ncell[axis]) {
    set_error("pga_fill_periodic: a halo %d deep over %d cells", depth, ncell[axis]);
    return PGA_EINVAL;
  }
  if (depth == 0) return PGA_OK;
  for (int a = 0; a < 3; a++)
    PGA_REQUIRE(ncell[a] > 0 && A.n[a] >= ncell[a] + 2 * ng, "pga_fill_periodic: the array is smaller than its cells and halos");
  PGA_REQUIRE(A.data && A.n[3] >= ncomp, "pga_fill_periodic: the array holds too few components");
  int ext[3] = {ncell[0], ncell[1], ncell[2]};
  ext[axis] = depth;
  if ((int64_t)ext[2] * ncomp > 65535) {
    set_error("pga_fill_periodic: too many k-planes times components");
    return PGA_ERANGE;
  }
  const dim3 grid((ext[0] + TI - 1) / TI, (ext[1] + TJ - 1) / TJ, ext[2] * ncomp);
  hipLaunchKernelGGL(k_fill_periodic, grid, dim3(TI, TJ), 0, stream_of(stream), view(A), ncomp, ncell[0], ncell[1],
                     ncell[2], ng, depth, axis);
  return launched("pga_fill_periodic");
}

extern "C" int32_t pga_source(const pga_case *c, pga_view Q, pga_view state, pga_view dQ, pga_box box, void *stream) {
  if (!c || !box_ok(box)) {
    set_error("pga_source: null case or bad box");
    return PGA_EINVAL;
  }
  if (c->desc.mechanism < 0 || !c->chem) {
    set_error("pga_source: the case has no compiled mechanism");
    return PGA_EUNSUP;
  }
  if (empty(box)) return PGA_OK;
  PGA_REQUIRE(covers(Q, box, c->desc.ns + 4) && covers(state, box, nstate(c->desc)) && covers(dQ, box, c->desc.ns + 4),
              "pga_source: the box reaches outside Q, state or dQ");
  const Box3 b = box3(box);
  bool done = false;
#define X(NAME, IDX)                                                                                       \
  if (c->desc.mechanism == IDX) {                                                                          \
    hipLaunchKernelGGL((k_source<Mech_##NAME>), grid_of(b), dim3(TI, TJ), 0, stream_of(stream), c->K, c->chem, \
                       view(Q), view(state), view(dQ), b);                                                 \
    done = true;                                                                                           \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_source: no mechanism %d", c->desc.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_source");
}

extern "C" int32_t pga_production(const pga_case *c, int32_t n, const double *rho, const double *Y, const double *T,
                                  double *omega, void *stream) {
  if (!c || n < 0 || (n && (!rho || !Y || !T || !omega))) {
    set_error("pga_production: bad arguments");
    return PGA_EINVAL;
  }
  if (c->desc.mechanism < 0 || !c->chem) {
    set_error("pga_production: the case has no compiled mechanism");
    return PGA_EUNSUP;
  }
  if (n == 0) return PGA_OK;
  bool done = false;
#define X(NAME, IDX)                                                                                          \
  if (c->desc.mechanism == IDX) {                                                                             \
    hipLaunchKernelGGL((k_production<Mech_##NAME>), dim3((n + 255) / 256), dim3(256), 0, stream_of(stream), c->K, \
                       c->chem, n, rho, Y, T, omega);                                                         \
    done = true;                                                                                              \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_production: no mechanism %d", c->desc.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_production");
}

// place.mako:11-37: a box of every component copied from src to dst, the source cells
// remapped by the interface's orientation
namespace pga {
struct CopyArgs {
  int d0, d1, d2, e0, e1, e2, ncomp;
  int64_t sbase, st0, st1, st2;
};

__global__ __launch_bounds__(256) void k_copy_box(View dst, View src, CopyArgs a) {
  const int ii = blockIdx.x * TI + threadIdx.x, jj = blockIdx.y * TJ + threadIdx.y;
  const int kk = blockIdx.z % a.e2, comp = blockIdx.z / a.e2;
  if (ii >= a.e0 || jj >= a.e1 || comp >= a.ncomp) return;
  const int64_t os = a.sbase + ii * a.st0 + jj * a.st1 + kk * a.st2;
  dst.at(dst.off(a.d0 + ii, a.d1 + jj, a.d2 + kk), comp) = src.at(os, comp);
}

// one cell of a box across a rotational periodic: the vector components [vec, vec+3) turned by R;
// with rows, the components are rows of three (each a gradient, turned), and the rows [vec, vec+3)
// (a velocity gradient) are turned down each column as well
__device__ __forceinline__ void copy_turned_cell(const View &dst, const View &src, int64_t od, int64_t os, int ncomp,
                                                 const Turn &t, int vec, int rows) {
  if (!rows) {
    for (int c = 0; c < ncomp; c++)
      if (c < vec || c >= vec + 3) dst.at(od, c) = src.at(os, c);
    if (vec >= 0)
      turned(t.R, src.at(os, vec), src.at(os, vec + 1), src.at(os, vec + 2), dst.at(od, vec), dst.at(od, vec + 1),
             dst.at(od, vec + 2));
    return;
  }
  for (int r = 0; r < ncomp / 3; r++)
    if (r < vec || r >= vec + 3)
      turned(t.R, src.at(os, 3 * r), src.at(os, 3 * r + 1), src.at(os, 3 * r + 2), dst.at(od, 3 * r),
             dst.at(od, 3 * r + 1), dst.at(od, 3 * r + 2));
  if (vec < 0) return;
  double H[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const int c = 3 * (vec + r);
    turned(t.R, src.at(os, c), src.at(os, c + 1), src.at(os, c + 2), H[r][0], H[r][1], H[r][2]);
  }
#pragma unroll
  for (int e = 0; e < 3; e++)
    turned(t.R, H[0][e], H[1][e], H[2][e], dst.at(od, 3 * vec + e), dst.at(od, 3 * (vec + 1) + e),
           dst.at(od, 3 * (vec + 2) + e));
}

// the same box across a rotational periodic, a cell per thread
__global__ __launch_bounds__(256) void k_copy_turned(View dst, View src, CopyArgs a, Turn t, int vec, int rows) {
  const int ii = blockIdx.x * TI + threadIdx.x, jj = blockIdx.y * TJ + threadIdx.y, kk = blockIdx.z;
  if (ii >= a.e0 || jj >= a.e1) return;
  const int64_t os = a.sbase + ii * a.st0 + jj * a.st1 + kk * a.st2, od = dst.off(a.d0 + ii, a.d1 + jj, a.d2 + kk);
  copy_turned_cell(dst, src, od, os, a.ncomp, t, vec, rows);
}

// up to six slabs in one launch, a cell per thread over the slabs back to back (i fastest in a
// slab), every component copied, or turned where the slab's turn is on. Each slab begins on a
// workgroup's boundary, so a workgroup's slab is a scalar choice read from the kernel arguments
// once, not a per-lane one that would read them through vector loads
struct SlabCopy {
  View dst, src;
  CopyArgs a;
  Turn t;
  int64_t first, cells;
};

struct SlabsArgs {
  SlabCopy s[6];
  int n, ncomp, vec, rows;
};

__global__ __launch_bounds__(256) void k_copy_slabs(SlabsArgs a) {
  const int64_t c0 = (int64_t)blockIdx.x * 256;
  int n = 0;
  while (n + 1 < a.n && c0 >= a.s[n + 1].first) n++;
  const SlabCopy &sl = a.s[n];
  // a slab holds fewer than 2^31 cells
  const unsigned r = (unsigned)(c0 - sl.first) + threadIdx.x;
  if (r >= sl.cells) return;
  const unsigned e0 = sl.a.e0, e1 = sl.a.e1;
  const int ii = (int)(r % e0), jj = (int)((r / e0) % e1), kk = (int)(r / (e0 * e1));
  const int64_t os = sl.a.sbase + ii * sl.a.st0 + jj * sl.a.st1 + kk * sl.a.st2;
  const int64_t od = sl.dst.off(sl.a.d0 + ii, sl.a.d1 + jj, sl.a.d2 + kk);
  if (sl.t.on) {
    copy_turned_cell(sl.dst, sl.src, od, os, a.ncomp, sl.t, a.vec, a.rows);
    return;
  }
  // eight components read before any is written, so a thread waits on memory once per eight
  for (int c0 = 0; c0 < a.ncomp; c0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (c0 + u < a.ncomp) v[u] = sl.src.at(os, c0 + u);
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (c0 + u < a.ncomp) sl.dst.at(od, c0 + u) = v[u];
  }
}
}  // namespace pga

extern "C" int32_t pga_copy_box(pga_view dst, const int32_t *dst_start, pga_view src, const int32_t *src_start,
                                const int32_t *extent, int32_t ncomp, const int32_t *orient, const double *rot,
                                int32_t vec, int32_t rows, void *stream) {
  if (!dst_start || !src_start || !extent || ncomp <= 0) {
    set_error("pga_copy_box: null argument or no components");
    return PGA_EINVAL;
  }
  for (int a = 0; a < 3; a++)
    if (dst_start[a] < 0 || src_start[a] < 0 || extent[a] < 0) {
      set_error("pga_copy_box: negative start or extent");
      return PGA_EINVAL;
    }
  if (!extent[0] || !extent[1] || !extent[2]) return PGA_OK;
  Placement pl;
  {
    const pga_box bd{{dst_start[0], dst_start[1], dst_start[2]}, {extent[0], extent[1], extent[2]}};
    PGA_REQUIRE(covers(dst, bd, ncomp), "pga_copy_box: the destination box reaches outside its array");
    if (!placement("pga_copy_box", src, src_start, orient, extent, ncomp, pl)) return PGA_ERANGE;
  }
  if ((int64_t)extent[2] * ncomp > 65535) {
    set_error("pga_copy_box: too many k-planes times components");
    return PGA_ERANGE;
  }
  const CopyArgs a{dst_start[0], dst_start[1], dst_start[2], extent[0], extent[1], extent[2], ncomp,
                   pl.base,      pl.step[0],   pl.step[1],   pl.step[2]};
  const Turn t = turn_of(rot);
  if (t.on) {
    const int units = rows ? ncomp / 3 : ncomp;
    PGA_REQUIRE((rows == 0 || rows == 1) && (!rows || ncomp % 3 == 0) && vec >= -1 && vec + 3 <= units,
                "pga_copy_box: a turned copy's vector (%d) must lie within its %d %s", vec, units,
                rows ? "rows" : "components");
    const dim3 tgrid((extent[0] + TI - 1) / TI, (extent[1] + TJ - 1) / TJ, extent[2]);
    hipLaunchKernelGGL(k_copy_turned, tgrid, dim3(TI, TJ), 0, stream_of(stream), view(dst), view(src), a, t, vec,
                       rows);
    return launched("pga_copy_box");
  }
  const dim3 grid((extent[0] + TI - 1) / TI, (extent[1] + TJ - 1) / TJ, extent[2] * ncomp);
  hipLaunchKernelGGL(k_copy_box, grid, dim3(TI, TJ), 0, stream_of(stream), view(dst), view(src), a);
  return launched("pga_copy_box");
}

extern "C" int32_t pga_copy_slabs(int32_t ntask, const pga_view *dst, const pga_view *src, const int32_t *dst_start,
                                  const int32_t *src_start, const int32_t *extent, int32_t ncomp, const int32_t *orient,
                                  const double *rot, int32_t vec, int32_t rows, void *stream) {
  if (ntask < 0 || ntask > 6 || ncomp <= 0 || (ntask && (!dst || !src || !dst_start || !src_start || !extent))) {
    set_error("pga_copy_slabs: 0 to 6 slabs with their views, starts and extents, and components");
    return PGA_EINVAL;
  }
  SlabsArgs a{};
  a.ncomp = ncomp, a.vec = vec, a.rows = rows;
  int64_t total = 0;
  for (int n = 0; n < ntask; n++) {
    const int32_t *ds = dst_start + 3 * n, *ss = src_start + 3 * n, *ex = extent + 3 * n;
    for (int m = 0; m < 3; m++)
      PGA_REQUIRE(ds[m] >= 0 && ss[m] >= 0 && ex[m] >= 0, "pga_copy_slabs: slab %d has a negative start or extent", n);
    if (!ex[0] || !ex[1] || !ex[2]) continue;
    const pga_box bd{{ds[0], ds[1], ds[2]}, {ex[0], ex[1], ex[2]}};
    PGA_REQUIRE(covers(dst[n], bd, ncomp), "pga_copy_slabs: slab %d's destination reaches outside its array", n);
    Placement pl;
    if (!placement("pga_copy_slabs", src[n], ss, orient ? orient + 3 * n : nullptr, ex, ncomp, pl)) return PGA_ERANGE;
    SlabCopy &sl = a.s[a.n++];
    sl.dst = view(dst[n]), sl.src = view(src[n]);
    sl.a = CopyArgs{ds[0], ds[1], ds[2], ex[0], ex[1], ex[2], ncomp, pl.base, pl.step[0], pl.step[1], pl.step[2]};
    sl.t = turn_of(rot ? rot + 9 * n : nullptr);
    if (sl.t.on) {
      const int units = rows ? ncomp / 3 : ncomp;
      PGA_REQUIRE((rows == 0 || rows == 1) && (!rows || ncomp % 3 == 0) && vec >= -1 && vec + 3 <= units,
                  "pga_copy_slabs: a turned copy's vector (%d) must lie within its %d %s", vec, units,
                  rows ? "rows" : "components");
    }
    sl.first = total;
    sl.cells = (int64_t)ex[0] * ex[1] * ex[2];
    total += (sl.cells + 255) / 256 * 256;
  }
  if (!total) return PGA_OK;
  PGA_REQUIRE(total < FLAT_MAX, "pga_copy_slabs: %lld cells; at most 2^31", (long long)total);
  hipLaunchKernelGGL(k_copy_slabs, dim3((unsigned)(total / 256)), dim3(256), 0, stream_of(stream), a);
  return launched("pga_copy_slabs");
}
