// Walls (simulator/euler/bcs.py:15-27, navstokes/bcs.py:27-53): each wall's halo from the interior's
// image across it (the preadvflux hook) and the wall's own flux through its faces (the postflux
// hook), for the six kinds the reference names:
//   slp-adia-wall, slp-isot-wall      the velocity mirrored across the wall (euler/kernels/bcs/
//                                     slp-{adia,isot}-wall.mako); no shear
//   no-slp-adia-wall, no-slp-isot-wall the velocity held at zero, and mov-adia-wall, mov-isot-wall
//   mov-adia-wall, mov-isot-wall      at the wall's own (navstokes/kernels/bcs/{no-slp,mov}-*.mako,
//                                     wall.mako stuckhalo, stuckhaloisot)
// an isothermal wall's halo takes the wall's T (its state by the EOS from p, T_w), and its flux the
// heat flux from T_w and the transport properties at the wall (wall.mako wallcoeffs, isotflux).
//
// One thread per (t1, t2) position on the wall: t1 = (axis + 1) % 3 runs along the lanes when it is
// i, so the mirror's loads along i coalesce for the j and k walls.
#include "pga_internal.hpp"
#include "pga_trans.hpp"

namespace pga {

enum { WALL_SLP_ADIA = 0, WALL_SLP_ISOT = 1, WALL_NOSLP_ADIA = 2, WALL_NOSLP_ISOT = 3, WALL_MOV_ADIA = 4,
       WALL_MOV_ISOT = 5 };
enum { TRANS_NONE = -1 };

__host__ __device__ inline bool wall_isot(int kind) { return kind == WALL_SLP_ISOT || kind == WALL_NOSLP_ISOT || kind == WALL_MOV_ISOT; }
__host__ __device__ inline bool wall_stuck(int kind) { return kind >= WALL_NOSLP_ADIA; }

struct WallArgs {
  const double *K;
  View Q, st, S, tr, xcc, xf, F;
  int axis, side, ng, n, n1, n2, ne, nstate, viscous, kind;
  double wall[3], Tw;
};

__device__ __forceinline__ void wall_index(const WallArgs &a, int plane, int b1, int b2, int (&ix)[3]) {
  const int t1 = (a.axis + 1) % 3, t2 = (a.axis + 2) % 3;
  ix[a.axis] = plane;
  ix[t1] = a.ng + b1;
  ix[t2] = a.ng + b2;
}

// solution.mako:14-23 fromprims: a cell's Q and state from p, T, the velocity and the carried Y
template <int NS, int EOS, class YF>
__device__ __forceinline__ void fromprims_cell(const Gas<NS> &g, double p, double T, const double (&u)[3], const YF &Y,
                                               const View &Q, const View &st, int64_t o) {
#pragma clang fp contract(off)
  MixMom<NS, EOS> m;
  mixmom<NS, EOS>(g, Y, m);
  EosState s;
  frompT<NS, EOS>(g, m, Y, p, T, s);
  Q.at(o, 0) = s.rho;
  Q.at(o, 1) = s.rho * u[0];
  Q.at(o, 2) = s.rho * u[1];
  Q.at(o, 3) = s.rho * u[2];
  Q.at(o, 4) = s.rho * (s.e + 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
#pragma unroll
  for (int n = 0; n < NS - 1; n++) Q.at(o, 5 + n) = s.rho * Y(n);
  store_state<EOS>(st, o, s);
}

// the halo cell h planes out from the interior cell h planes in
template <int NS, int EOS>
__global__ __launch_bounds__(256) void k_wall_mirror(WallArgs a) {
#pragma clang fp contract(off)
  const int b1 = blockIdx.x * 64 + (threadIdx.x & 63), b2 = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (b1 >= a.n1 || b2 >= a.n2) return;
  const Gas<NS> g{a.K};
  int f[3];
  wall_index(a, a.side ? a.ng + a.n : a.ng, b1, b2, f);
  const int64_t of = a.S.off(f[0], f[1], f[2]);
  const double Sx = a.S.at(of, 0), Sy = a.S.at(of, 1), Sz = a.S.at(of, 2);
  const double Sinv = 1.0 / sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
  const double nx = Sx * Sinv, ny = Sy * Sinv, nz = Sz * Sinv;
  for (int h = 0; h < a.ng; h++) {
    // a block thinner than the halo reflects its last cell into every deeper layer
    // (launch/policy.py:168-184, MirrorOffset)
    const int hm = h < a.n ? h : a.n - 1;
    int ih[3], im[3];
    wall_index(a, a.side ? a.ng + a.n + h : a.ng - 1 - h, b1, b2, ih);
    wall_index(a, a.side ? a.ng + a.n - 1 - hm : a.ng + hm, b1, b2, im);
    const int64_t oh = a.Q.off(ih[0], ih[1], ih[2]), om = a.Q.off(im[0], im[1], im[2]);
    const int64_t sh = a.st.off(ih[0], ih[1], ih[2]), sm = a.st.off(im[0], im[1], im[2]);
    if (a.kind == WALL_SLP_ADIA) {
      // the momentum's normal part, taken off twice; the whole state the image's (no EOS runs)
      const double q1 = a.Q.at(om, 1), q2 = a.Q.at(om, 2), q3 = a.Q.at(om, 3);
      const double mn = 2.0 * (q1 * nx + q2 * ny + q3 * nz);
      a.Q.at(oh, 0) = a.Q.at(om, 0);
      a.Q.at(oh, 1) = q1 - mn * nx;
      a.Q.at(oh, 2) = q2 - mn * ny;
      a.Q.at(oh, 3) = q3 - mn * nz;
      for (int l = 4; l < a.ne; l++) a.Q.at(oh, l) = a.Q.at(om, l);
      for (int c = 0; c < a.nstate; c++) a.st.at(sh, c) = a.st.at(sm, c);
    } else if (a.kind == WALL_NOSLP_ADIA || a.kind == WALL_MOV_ADIA) {
      // stuckhalo: the velocity reflected about the wall's, the energy carrying the change in the
      // kinetic part, the state the image's
      const double rho = a.Q.at(om, 0), rhoinv = 1.0 / rho;
      double ke = 0.0, kehalo = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double u = a.Q.at(om, 1 + d) * rhoinv, uh = 2.0 * a.wall[d] - u;
        a.Q.at(oh, 1 + d) = rho * uh;
        ke += u * u;
        kehalo += uh * uh;
      }
      a.Q.at(oh, 0) = rho;
      a.Q.at(oh, 4) = a.Q.at(om, 4) + 0.5 * rho * (kehalo - ke);
      for (int l = 5; l < a.ne; l++) a.Q.at(oh, l) = a.Q.at(om, l);
      for (int c = 0; c < a.nstate; c++) a.st.at(sh, c) = a.st.at(sm, c);
    } else {
      // an isothermal wall: the image's p and mass fractions at the wall's T, by the EOS
      const double rhoinv = 1.0 / a.Q.at(om, 0);
      double u[3];
      if (a.kind == WALL_SLP_ISOT) {
        const double uin = a.Q.at(om, 1) * rhoinv, vin = a.Q.at(om, 2) * rhoinv, win = a.Q.at(om, 3) * rhoinv;
        const double un = 2.0 * (uin * nx + vin * ny + win * nz);
        u[0] = uin - un * nx;
        u[1] = vin - un * ny;
        u[2] = win - un * nz;
      } else {
#pragma unroll
        for (int d = 0; d < 3; d++) u[d] = 2.0 * a.wall[d] - a.Q.at(om, 1 + d) * rhoinv;
      }
      double Y[NS > 1 ? NS - 1 : 1];
#pragma unroll
      for (int n = 0; n < NS - 1; n++) Y[n] = a.Q.at(om, 5 + n) * rhoinv;
      fromprims_cell<NS, EOS>(g, a.st.at(sm, 0), a.Tw, u, [&](int n) { return Y[n]; }, a.Q, a.st, oh);
    }
  }
}

// the wall's flux from the cell next to it: p_w = max(p + rho c u_n, 0), u_n toward the wall
// (wall.mako wallpressure); a viscous wall adds the stress of the two-point normal velocity gradient
// relative to the wall (every component where the flow is held, the normal one where it slips), the
// wall's work, and the heat flux from T_w (navstokes wall.mako wallfaceflux)
template <int NS, int EOS, int TRANS>
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
  if constexpr (TRANS != TRANS_NONE) {
    const bool isot = wall_isot(a.kind);
    const double stuck = wall_stuck(a.kind) ? 1.0 : 0.0;
    double mu, kappa, dT;
    if (isot) {
      // wallcoeffs: the transport properties at the wall, the cell's composition at p and T_w
      const Gas<NS> g{a.K};
      const double rhoinv = 1.0 / q[0];
      double Y[NS], Ylast = 1.0;
#pragma unroll
      for (int n = 0; n < NS - 1; n++) {
        Y[n] = a.Q.at(oc, 5 + n) * rhoinv;
        Ylast -= a.Q.at(oc, 5 + n) * rhoinv;
      }
      Y[NS - 1] = fmax(Ylast, 0.0);
      MixMom<NS, EOS> m;
      mixmom<NS, EOS>(g, [&](int n) { return Y[n]; }, m);
      EosState thermo;
      frompT<NS, EOS>(g, m, [&](int n) { return Y[n]; }, a.st.at(osc, 0), a.Tw, thermo);
      double denom = 0.0;
#pragma unroll
      for (int n = 0; n < NS; n++) denom += Y[n] * g.MWinv(n);
      const double MWmix = 1.0 / denom;
      double X[NS];
#pragma unroll
      for (int n = 0; n < NS; n++) X[n] = Y[n] * g.MWinv(n) * MWmix;
      mu_kappa<NS, TRANS>(g, X, a.Tw, elog(a.Tw), thermo.rho, mu, kappa);
      dT = a.Tw - a.st.at(osc, 1);
    } else {
      const int64_t ot = a.tr.off(ic[0], ic[1], ic[2]);
      mu = a.tr.at(ot, 0);
      kappa = a.tr.at(ot, 1);
      dT = 0.0;
    }
    const int64_t ox = a.xcc.off(ic[0], ic[1], ic[2]), oxf = a.xf.off(f[0], f[1], f[2]);
    const double Sabs = sqrt(S3[0] * S3[0] + S3[1] * S3[1] + S3[2] * S3[2]), Sinv = 1.0 / Sabs;
    const double nv[3] = {S3[0] * Sinv, S3[1] * Sinv, S3[2] * Sinv};
    const double dn = fabs((a.xcc.at(ox, 0) - a.xf.at(oxf, 0)) * nv[0] + (a.xcc.at(ox, 1) - a.xf.at(oxf, 1)) * nv[1] +
                           (a.xcc.at(ox, 2) - a.xf.at(oxf, 2)) * nv[2]);
    const double rhoinv = 1.0 / q[0];
    double dun[3];
    for (int d = 0; d < 3; d++) dun[d] = outward * (a.wall[d] - q[1 + d] * rhoinv) / dn;
    const double div = dun[0] * nv[0] + dun[1] * nv[1] + dun[2] * nv[2];
    for (int d = 0; d < 3; d++) dun[d] = div * nv[d] + stuck * (dun[d] - div * nv[d]);
    double work = 0.0, tS[3];
    for (int d = 0; d < 3; d++) {
      tS[d] = -mu * Sabs * (dun[d] + div * nv[d] / 3.0);
      work += a.wall[d] * tS[d];
    }
    F1 += tS[0];
    F2 += tS[1];
    F3 += tS[2];
    F4 += work - kappa * outward * (dT) / dn * Sabs;
  }
  const int64_t off_ = a.F.off(f[0], f[1], f[2]);
  a.F.at(off_, 0) = 0.0;
  a.F.at(off_, 1) = F1;
  a.F.at(off_, 2) = F2;
  a.F.at(off_, 3) = F3;
  a.F.at(off_, 4) = F4;
  for (int l = 5; l < a.ne; l++) a.F.at(off_, l) = 0.0;
}

static bool wall_args(const pga_case *c, int axis, int side, const pga_wall *w, const int32_t *ncell, int ng,
                      WallArgs &a, const char *who) {
  if (!c || axis < 0 || axis > 2 || side < 0 || side > 1 || !ncell || ng < 1) {
    set_error("%s: null case or ncell, an axis other than 0..2, a side other than 0 or 1, or ng < 1", who);
    return false;
  }
  for (int d = 0; d < 3; d++)
    if (ncell[d] < 1) {
      set_error("%s: a block of %d x %d x %d cells", who, ncell[0], ncell[1], ncell[2]);
      return false;
    }
  a.K = c->K;
  a.axis = axis, a.side = side, a.ng = ng, a.n = ncell[axis];
  a.n1 = ncell[(axis + 1) % 3], a.n2 = ncell[(axis + 2) % 3];
  a.ne = c->desc.ns + 4;
  a.nstate = nstate(c->desc);
  a.kind = w ? w->kind : WALL_SLP_ADIA;
  a.Tw = w ? w->T : 0.0;
  for (int d = 0; d < 3; d++) a.wall[d] = w && a.kind >= WALL_MOV_ADIA ? w->u[d] : 0.0;
  if (a.kind < WALL_SLP_ADIA || a.kind > WALL_MOV_ISOT) {
    set_error("%s: no wall kind %d (0 slp-adia, 1 slp-isot, 2 no-slp-adia, 3 no-slp-isot, 4 mov-adia, 5 mov-isot)",
              who, a.kind);
    return false;
  }
  if (wall_isot(a.kind) && !(a.Tw > 0.0)) {
    set_error("%s: an isothermal wall at T = %g", who, a.Tw);
    return false;
  }
  if (a.kind >= WALL_NOSLP_ADIA && c->desc.trans < 0) {
    set_error("%s: a wall that holds the flow needs a viscous case (an Euler case has slp-adia and slp-isot walls)", who);
    return false;
  }
  return true;
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_wall_mirror(const pga_case *c, int32_t axis, int32_t side, const pga_wall *w,
                                   const int32_t *ncell, int32_t ng, pga_view Q, pga_view state, pga_view S,
                                   void *stream) {
  WallArgs a{};
  if (!wall_args(c, axis, side, w, ncell, ng, a, "pga_wall_mirror")) return PGA_EINVAL;
  // the halo slab ng deep and its mirror: the whole padded extent along the axis, the interior across
  pga_box cells{{ng, ng, ng}, {ncell[0], ncell[1], ncell[2]}};
  cells.start[axis] = 0;
  cells.extent[axis] = ncell[axis] + 2 * ng;
  pga_box faces{{ng, ng, ng}, {ncell[0], ncell[1], ncell[2]}};
  faces.start[axis] = side ? ng + ncell[axis] : ng;
  faces.extent[axis] = 1;
  PGA_REQUIRE(covers(Q, cells, a.ne) && covers(state, cells, a.nstate) && covers(S, faces, 3),
              "pga_wall_mirror: the halo, its mirror or the wall faces reach outside Q, state or S");
  PGA_REQUIRE(Q.sj == state.sj && Q.sk == state.sk, "pga_wall_mirror: Q and state must share their layout");
  a.Q = view(Q), a.st = view(state), a.S = view(S);
  const dim3 grid((a.n1 + 63) / 64, (a.n2 + 3) / 4);
  const bool done = dispatch_eos(c->desc, [&]<int N, int E>() {
    hipLaunchKernelGGL((k_wall_mirror<N, E>), grid, dim3(256), 0, stream_of(stream), a);
  });
  if (!done) {
    set_error("pga_wall_mirror: no kernel for %d species with eos %d", c->desc.ns, c->desc.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_wall_mirror");
}

extern "C" int32_t pga_wall_flux(const pga_case *c, int32_t axis, int32_t side, const pga_wall *w,
                                 const int32_t *ncell, int32_t ng, pga_view Q, pga_view state, pga_view trstate,
                                 pga_view xcc, pga_view xf, pga_view S, pga_view F, void *stream) {
  WallArgs a{};
  if (!wall_args(c, axis, side, w, ncell, ng, a, "pga_wall_flux")) return PGA_EINVAL;
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
  PGA_REQUIRE(a.viscous || a.kind <= WALL_SLP_ISOT, "pga_wall_flux: a wall that holds the flow needs trstate");
  a.Q = view(Q), a.st = view(state), a.S = view(S), a.F = view(F);
  if (a.viscous) a.tr = view(trstate), a.xcc = view(xcc), a.xf = view(xf);
  const dim3 grid((a.n1 + 63) / 64, (a.n2 + 3) / 4);
  const int trans = a.viscous ? c->desc.trans : TRANS_NONE;
  const bool done = dispatch_eos(c->desc, [&]<int N, int E>() {
    if (trans == TRANS_NONE)
      hipLaunchKernelGGL((k_wall_flux<N, E, TRANS_NONE>), grid, dim3(256), 0, stream_of(stream), a);
    else if (trans == TRANS_CONSTANT)
      hipLaunchKernelGGL((k_wall_flux<N, E, TRANS_CONSTANT>), grid, dim3(256), 0, stream_of(stream), a);
    else if (trans == TRANS_KINETIC)
      hipLaunchKernelGGL((k_wall_flux<N, E, TRANS_KINETIC>), grid, dim3(256), 0, stream_of(stream), a);
    else if constexpr (E == EOS_REAL)
      hipLaunchKernelGGL((k_wall_flux<N, E, TRANS_CHUNG>), grid, dim3(256), 0, stream_of(stream), a);
  });
  if (!done) {
    set_error("pga_wall_flux: no kernel for %d species with eos %d", c->desc.ns, c->desc.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_wall_flux");
}
