// The adiabatic slip wall (simulator/euler/bcs.py:20-27, navstokes/bcs.py:27-53): the halo as the
// interior's mirror image across the wall (the preadvflux hook, euler/kernels/bcs/slp-adia-wall.mako)
// and the wall's own flux through its faces (the postflux hook, euler/kernels/bcs/wall.mako and
// navstokes/kernels/bcs/{slp-adia-wall,wall}.mako).
//
// One thread per (t1, t2) position on the wall: t1 = (axis + 1) % 3 runs along the lanes when it is
// i, so the mirror's loads along i coalesce for the j and k walls.
#include "pga_common.hpp"
#include "pga_internal.hpp"

namespace pga {

struct WallArgs {
  View Q, st, S, tr, xcc, xf, F;
  int axis, side, ng, n, n1, n2, ne, nstate, viscous;
};

__device__ __forceinline__ void wall_index(const WallArgs &a, int plane, int b1, int b2, int (&ix)[3]) {
  const int t1 = (a.axis + 1) % 3, t2 = (a.axis + 2) % 3;
  ix[a.axis] = plane;
  ix[t1] = a.ng + b1;
  ix[t2] = a.ng + b2;
}

// the halo cell h planes out takes the interior cell h planes in, its momentum's normal part
// (along the wall face's unit normal at that position) taken off twice, its state copied
__global__ __launch_bounds__(256) void k_wall_mirror(WallArgs a) {
#pragma clang fp contract(off)
  const int b1 = blockIdx.x * 64 + (threadIdx.x & 63), b2 = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (b1 >= a.n1 || b2 >= a.n2) return;
  int f[3];
  wall_index(a, a.side ? a.ng + a.n : a.ng, b1, b2, f);
  const int64_t of = a.S.off(f[0], f[1], f[2]);
  const double Sx = a.S.at(of, 0), Sy = a.S.at(of, 1), Sz = a.S.at(of, 2);
  const double Sinv = 1.0 / sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
  const double nx = Sx * Sinv, ny = Sy * Sinv, nz = Sz * Sinv;
  for (int h = 0; h < a.ng; h++) {
    int ih[3], im[3];
    wall_index(a, a.side ? a.ng + a.n + h : a.ng - 1 - h, b1, b2, ih);
    wall_index(a, a.side ? a.ng + a.n - 1 - h : a.ng + h, b1, b2, im);
    const int64_t oh = a.Q.off(ih[0], ih[1], ih[2]), om = a.Q.off(im[0], im[1], im[2]);
    const double q1 = a.Q.at(om, 1), q2 = a.Q.at(om, 2), q3 = a.Q.at(om, 3);
    const double mn = 2.0 * (q1 * nx + q2 * ny + q3 * nz);
    a.Q.at(oh, 0) = a.Q.at(om, 0);
    a.Q.at(oh, 1) = q1 - mn * nx;
    a.Q.at(oh, 2) = q2 - mn * ny;
    a.Q.at(oh, 3) = q3 - mn * nz;
    for (int l = 4; l < a.ne; l++) a.Q.at(oh, l) = a.Q.at(om, l);
    const int64_t sh = a.st.off(ih[0], ih[1], ih[2]), sm = a.st.off(im[0], im[1], im[2]);
    for (int c = 0; c < a.nstate; c++) a.st.at(sh, c) = a.st.at(sm, c);
  }
}

// the wall's flux from the cell next to it: p_w = max(p + rho c u_n, 0), u_n toward the wall
// (wall.mako wallpressure); a viscous case adds the normal stress of the two-point normal velocity
// gradient and no heat (navstokes wall.mako wallfaceflux with the flow slipping)
__global__ __launch_bounds__(256) void k_wall_flux(WallArgs a) {
#pragma clang fp contract(off)
  const int b1 = blockIdx.x * 64 + (threadIdx.x & 63), b2 = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (b1 >= a.n1 || b2 >= a.n2) return;
  int f[3], ic[3];
  wall_index(a, a.side ? a.ng + a.n : a.ng, b1, b2, f);
  wall_index(a, a.side ? a.ng + a.n - 1 : a.ng, b1, b2, ic);
  const double outward = a.side ? 1.0 : -1.0;
  const int64_t of = a.S.off(f[0], f[1], f[2]), oc = a.Q.off(ic[0], ic[1], ic[2]), osc = a.st.off(ic[0], ic[1], ic[2]);
  const double S3[3] = {a.S.at(of, 0), a.S.at(of, 1), a.S.at(of, 2)};
  const double q[4] = {a.Q.at(oc, 0), a.Q.at(oc, 1), a.Q.at(oc, 2), a.Q.at(oc, 3)};
  const double area = sqrt(S3[0] * S3[0] + S3[1] * S3[1] + S3[2] * S3[2]);
  const double un = outward * (q[1] * S3[0] + q[2] * S3[1] + q[3] * S3[2]) / (q[0] * area);
  const double pw = fmax(a.st.at(osc, 0) + q[0] * a.st.at(osc, 7) * un, 0.0);
  double F1 = pw * S3[0], F2 = pw * S3[1], F3 = pw * S3[2], F4 = 0.0;
  if (a.viscous) {
    const int64_t ot = a.tr.off(ic[0], ic[1], ic[2]), ox = a.xcc.off(ic[0], ic[1], ic[2]), oxf = a.xf.off(f[0], f[1], f[2]);
    const double mu = a.tr.at(ot, 0);
    const double Sabs = sqrt(S3[0] * S3[0] + S3[1] * S3[1] + S3[2] * S3[2]), Sinv = 1.0 / Sabs;
    const double nv[3] = {S3[0] * Sinv, S3[1] * Sinv, S3[2] * Sinv};
    const double dn = fabs((a.xcc.at(ox, 0) - a.xf.at(oxf, 0)) * nv[0] + (a.xcc.at(ox, 1) - a.xf.at(oxf, 1)) * nv[1] +
                           (a.xcc.at(ox, 2) - a.xf.at(oxf, 2)) * nv[2]);
    const double rhoinv = 1.0 / q[0];
    double dun[3];
    for (int d = 0; d < 3; d++) dun[d] = outward * (0.0 - q[1 + d] * rhoinv) / dn;
    const double div = dun[0] * nv[0] + dun[1] * nv[1] + dun[2] * nv[2];
    for (int d = 0; d < 3; d++) dun[d] = div * nv[d] + 0.0 * (dun[d] - div * nv[d]);
    double work = 0.0, tS[3];
    for (int d = 0; d < 3; d++) {
      tS[d] = -mu * Sabs * (dun[d] + div * nv[d] / 3.0);
      work += 0.0 * tS[d];
    }
    F1 += tS[0];
    F2 += tS[1];
    F3 += tS[2];
    F4 += work - a.tr.at(ot, 1) * outward * (0.0) / dn * Sabs;
  }
  const int64_t off_ = a.F.off(f[0], f[1], f[2]);
  a.F.at(off_, 0) = 0.0;
  a.F.at(off_, 1) = F1;
  a.F.at(off_, 2) = F2;
  a.F.at(off_, 3) = F3;
  a.F.at(off_, 4) = F4;
  for (int l = 5; l < a.ne; l++) a.F.at(off_, l) = 0.0;
}

static bool wall_args(const pga_case *c, int axis, int side, const int32_t *ncell, int ng, WallArgs &a, const char *who) {
  if (!c || axis < 0 || axis > 2 || side < 0 || side > 1 || !ncell || ng < 1) {
    set_error("%s: null case or ncell, an axis other than 0..2, a side other than 0 or 1, or ng < 1", who);
    return false;
  }
  for (int d = 0; d < 3; d++)
    if (ncell[d] < 1) {
      set_error("%s: a block of %d x %d x %d cells", who, ncell[0], ncell[1], ncell[2]);
      return false;
    }
  a.axis = axis, a.side = side, a.ng = ng, a.n = ncell[axis];
  a.n1 = ncell[(axis + 1) % 3], a.n2 = ncell[(axis + 2) % 3];
  a.ne = c->desc.ns + 4;
  a.nstate = nstate(c->desc);
  return true;
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_wall_mirror(const pga_case *c, int32_t axis, int32_t side, const int32_t *ncell, int32_t ng,
                                   pga_view Q, pga_view state, pga_view S, void *stream) {
  WallArgs a{};
  if (!wall_args(c, axis, side, ncell, ng, a, "pga_wall_mirror")) return PGA_EINVAL;
  // the halo slab ng deep and its mirror: the whole padded extent along the axis, the interior across
  pga_box cells{{ng, ng, ng}, {ncell[0], ncell[1], ncell[2]}};
  cells.start[axis] = 0;
  cells.extent[axis] = ncell[axis] + 2 * ng;
  pga_box faces{{ng, ng, ng}, {ncell[0], ncell[1], ncell[2]}};
  faces.start[axis] = side ? ng + ncell[axis] : ng;
  faces.extent[axis] = 1;
  PGA_REQUIRE(ncell[axis] >= ng, "pga_wall_mirror: a wall needs at least ng = %d interior cells to mirror", ng);
  PGA_REQUIRE(covers(Q, cells, a.ne) && covers(state, cells, a.nstate) && covers(S, faces, 3),
              "pga_wall_mirror: the halo, its mirror or the wall faces reach outside Q, state or S");
  PGA_REQUIRE(Q.sj == state.sj && Q.sk == state.sk, "pga_wall_mirror: Q and state must share their layout");
  a.Q = view(Q), a.st = view(state), a.S = view(S);
  const dim3 grid((a.n1 + 63) / 64, (a.n2 + 3) / 4);
  hipLaunchKernelGGL(k_wall_mirror, grid, dim3(256), 0, stream_of(stream), a);
  return launched("pga_wall_mirror");
}

extern "C" int32_t pga_wall_flux(const pga_case *c, int32_t axis, int32_t side, const int32_t *ncell, int32_t ng,
                                 pga_view Q, pga_view state, pga_view trstate, pga_view xcc, pga_view xf, pga_view S,
                                 pga_view F, void *stream) {
  WallArgs a{};
  if (!wall_args(c, axis, side, ncell, ng, a, "pga_wall_flux")) return PGA_EINVAL;
  pga_box cell{{ng, ng, ng}, {ncell[0], ncell[1], ncell[2]}};
  cell.start[axis] = side ? ng + ncell[axis] - 1 : ng;
  cell.extent[axis] = 1;
  pga_box faces = cell;
  faces.start[axis] = side ? ng + ncell[axis] : ng;
  a.viscous = c->desc.trans >= 0 && trstate.data != nullptr;
  PGA_REQUIRE(covers(Q, cell, a.ne) && covers(state, cell, a.nstate) && covers(S, faces, 3) && covers(F, faces, a.ne),
              "pga_wall_flux: the wall's cells or faces reach outside Q, state, S or F");
  if (a.viscous)
    PGA_REQUIRE(covers(trstate, cell, 2 + c->desc.ns) && covers(xcc, cell, 3) && covers(xf, faces, 3),
                "pga_wall_flux: the wall's cells or faces reach outside trstate, xcc or xf");
  a.Q = view(Q), a.st = view(state), a.S = view(S), a.F = view(F);
  if (a.viscous) a.tr = view(trstate), a.xcc = view(xcc), a.xf = view(xf);
  const dim3 grid((a.n1 + 63) / 64, (a.n2 + 3) / 4);
  hipLaunchKernelGGL(k_wall_flux, grid, dim3(256), 0, stream_of(stream), a);
  return launched("pga_wall_flux");
}
