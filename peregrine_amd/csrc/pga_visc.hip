// The viscous terms of a Navier-Stokes stage: the transport state and species enthalpies of
// each cell, the cell gradients of u, v, w, T and the carried mass fractions, and the
// alpha-damped viscous face flux added to the advective one.
//
// Layout: trstate [mu, kappa, D_0 .. D_{ns-1}], hsp [h_0 .. h_{ns-1}] (J/kg), grads
// [3 * l + e] = d(row l)/dx_e for the rows u, v, w, T, Y_0 .. Y_{ns-2}, Jinv [3 * r + e] =
// d xi_r / d x_e, xcc and the face centres [x, y, z].
//
// Arithmetic follows the oracle operation for operation (contraction off, ln T by elog), so a
// face's flux is the oracle's to the bit.
#include "pga_internal.hpp"
#include "pga_diff.hpp"
#include "pga_trans.hpp"

namespace pga {

// eos/realgas/kernels/eos.mako:192-273: the cubic's partial enthalpies per mass at the state's own
// root Z, -Ru T^2 d ln phi_n / dT plus the ideal part and the translation's -p c_n
template <int NS>
__device__ __forceinline__ void hpartial(const Gas<NS> &g, const double (&Y)[NS], double p, double T, double Z,
                                         const View &hsp, int64_t oh) {
#pragma clang fp contract(off)
  const double sig = g.cubic(2), eps = g.cubic(3);
  const auto YF = [&](int n) { return Y[n]; };
  MixMom<NS, EOS_REAL> m;
  mixmom<NS, EOS_REAL>(g, YF, m);
  Cubic c;
  cubicmix(m, c);
  attraction(g, c, YF, T);
  dimless(g, c, p, T);
  c.Z = Z;
  rootslope(c);
  double hdep, cpdep;
  departures(g, c, T, hdep, cpdep);
  const double lnT = elog(T), Tinv = Ru * c.invRuT, sqrtT = sqrt(T);
  const double ZT = c.dZdT, Bs = c.Bstar, BT = -Bs * Tinv;
  const double invZsB = 1.0 / (Z + sig * Bs), invZeB = 1.0 / (Z + eps * Bs);
  const double Iz = -c.logratio / (sig - eps);
  const double IzT = ((ZT + sig * BT) * invZsB - (ZT + eps * BT) * invZeB) / (sig - eps);
  const double dlnZmB = (ZT - BT) / (Z - Bs);
  const double invS1 = 1.0 / c.S1, invBm = 1.0 / c.bm, invamsqrtT = invS1 * invS1 / sqrtT;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    double hRT = 0.0;
#pragma unroll
    for (int mm = NT - 1; mm >= 0; mm--) hRT = fma(hRT, lnT, g.hpoly(n, mm));
    hRT += g.h298(n) * Tinv;
    const double qT = (c.qd - c.q) * Tinv;
    const double sqrtan = g.cubicsqrtac(n) * modulus(g, n, sqrtT);
    const double r = 2.0 * sqrtan * invS1;
    const double rT = (sqrtan * c.S2 - g.cubicda(n) * c.S1) * invamsqrtT;
    const double bb = g.cubicb(n) * invBm;
    const double qbar = c.q * (r - bb), qbarT = qT * (r - bb) + c.q * rT;
    const double dlnphidT = bb * c.dZdT - dlnZmB - qbarT * Iz - qbar * IzT;
    hsp.at(oh, n) = (hRT * T * Ru - Ru * T * T * dlnphidT - g.vshift(n) * p) * g.MWinv(n);
  }
}

// X[m] for a partner index m the table holds (uniform across lanes): a select over the species,
// which keeps the mole fractions in registers where a dynamic index would move them to memory
template <int NS>
__device__ __forceinline__ double pick(const double (&X)[NS], int m) {
  double r = X[0];
#pragma unroll
  for (int i = 1; i < NS; i++) r = m == i ? X[i] : r;
  return r;
}

// diffusion/kernels/binary.mako:11-62: each species' mixture-averaged coefficient from its binary
// pairs with the bath, fits of T^3/2 / D_nm at unit pressure divided by p; a species alone (its
// mole fraction within 64 eps of one, or none of a partial bath's others present) takes its
// self-diffusion coefficient
template <int NS>
__device__ __forceinline__ void binary_diffusion(const Gas<NS> &g, const double (&X)[NS], double p, double T,
                                                 double lnT, double (&D)[NS]) {
#pragma clang fp contract(off)
  const double invT32 = 1.0 / (T * sqrt(T));
  const int width = (int)g.diffmeta(0);
  const bool partial = g.diffmeta(1) != 0.0;
  const double tiny = 64.0 * 2.220446049250313e-16;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    double sum1 = 0.0, sum2 = 0.0, rest = 0.0, bathx = 0.0;
#pragma unroll
    for (int k = 0; k < NS; k++) {
      if (k < width) {
        // the fraction clipped at zero: a negative trace would let rest cancel while sum1 stays
        // tiny and positive
        int m;
        double Xm;
        if (partial) {
          m = (int)g.bathof(n, k);
          Xm = m != n ? fmax(pick(X, m), 0.0) : 0.0;
        } else {
          m = k + (k >= n);
          Xm = fmax(X[m < NS ? m : NS - 1], 0.0);
        }
        double Dinv = 0.0;
#pragma unroll
        for (int t = NT - 1; t >= 0; t--) Dinv = Dinv * lnT + g.pairfit(n, k, t);
        Dinv *= invT32;
        const double MWm = g.MW(m);
        sum1 += Xm * Dinv;
        sum2 += Xm * MWm * Dinv;
        rest += Xm * MWm;
        bathx += Xm;
      }
    }
    double Dinvself = 0.0;
#pragma unroll
    for (int t = NT - 1; t >= 0; t--) Dinvself = Dinvself * lnT + g.selffit(n, t);
    Dinvself *= invT32;
    const double Xn = fmax(X[n], 0.0);
    bool alone;
    double num, den;
    if (partial) {
      alone = 1.0 - Xn <= tiny || bathx <= tiny;
      num = alone ? 1.0 : rest * bathx;
      den = alone ? p * Dinvself : p * ((1.0 - Xn) * rest * sum1 + Xn * bathx * sum2);
    } else {
      alone = 1.0 - Xn <= tiny;
      num = alone ? 1.0 : rest;
      den = alone ? p * Dinvself : p * (rest * sum1 + Xn * sum2);
    }
    D[n] = (alone || (rest > 0.0 && den > 0.0)) ? num / den : 0.0;
  }
}

// navstokes/kernels/transport.mako:5-35: mu and kappa by Wilke's rule and the conductivity mean
// (mixrule/kernels/wilke.mako:3-17, transport/kernels/mixing.mako:3-10) from constant species
// values (transport/kernels/constant.mako:5-16) or the kinetic-theory fits
// (transport/kernels/dilute.mako:5-26); D_n by the Lewis numbers (diffusion/kernels/lewis.mako:3-7)
// or the binary pairs;
// h_n ideal (eos/kernels/species.mako:29-36)
template <int NS, int TRANS, int EOS>
__device__ __forceinline__ void transport_cell(const Gas<NS> &g, const View &Q, const View &st, int64_t o,
                                               const View &tr, const View &hsp, int64_t ot, int64_t oh,
                                               bool binary) {
#pragma clang fp contract(off)
  const double rho = Q.at(o, 0), rhoinv = 1.0 / rho;
  double Y[NS];
  double Ylast = 1.0;
#pragma unroll
  for (int n = 0; n < NS - 1; n++) {
    const double q = Q.at(o, 5 + n);
    Y[n] = q * rhoinv;
    Ylast -= q * rhoinv;
  }
  Y[NS - 1] = fmax(Ylast, 0.0);
  double denom = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) denom += Y[n] * g.MWinv(n);
  const double MWmix = 1.0 / denom;
  double X[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) X[n] = Y[n] * g.MWinv(n) * MWmix;
  const double T = st.at(o, 1), cp = st.at(o, 5), lnT = elog(T);
  double mu, kappa;
  mu_kappa<NS, TRANS>(g, X, T, lnT, rho, mu, kappa);
  tr.at(ot, 0) = mu;
  tr.at(ot, 1) = kappa;
  if (binary) {
    double D[NS];
    binary_diffusion(g, X, st.at(o, 0), T, lnT, D);
#pragma unroll
    for (int n = 0; n < NS; n++) tr.at(ot, 2 + n) = D[n];
  } else {
    const double kappaoverrhocp = kappa * rhoinv / cp;
#pragma unroll
    for (int n = 0; n < NS; n++) tr.at(ot, 2 + n) = kappaoverrhocp / g.lewis(n);
  }
  if constexpr (EOS == EOS_REAL) {
    hpartial(g, Y, st.at(o, 0), T, st.at(o, 8), hsp, oh);
  } else {
    const double Tinv = 1.0 / T;
#pragma unroll
    for (int n = 0; n < NS; n++) {
      double hRT = 0.0;
#pragma unroll
      for (int m = NT - 1; m >= 0; m--) hRT = fma(hRT, lnT, g.hpoly(n, m));
      hRT += g.h298(n) * Tinv;
      hsp.at(oh, n) = hRT * T * Ru * g.MWinv(n);
    }
  }
}

// the boxes of one transport call back to back, a cell per thread (i fastest in a box)
struct BoxList {
  Box3 b[8];
  int64_t first[8];
  int n;
};

// the transport of the cell the thread takes in a box list (a workgroup's box a scalar choice, each
// box beginning on a workgroup's boundary); false past its box
__device__ __forceinline__ bool boxlist_cell(const BoxList &L, unsigned blk, int &i, int &j, int &k) {
  const int64_t c0 = (int64_t)blk * 256;
  int n = 0;
  while (n + 1 < L.n && c0 >= L.first[n + 1]) n++;
  const Box3 &b = L.b[n];
  // fewer than 2^31 cells in all
  const unsigned r = (unsigned)(c0 - L.first[n]) + threadIdx.x;
  const unsigned e0 = b.e0, e1 = b.e1;
  if (r >= e0 * e1 * (unsigned)b.e2) return false;
  i = b.s0 + (int)(r % e0), j = b.s1 + (int)((r / e0) % e1), k = b.s2 + (int)(r / (e0 * e1));
  return true;
}

template <int NS, int TRANS, int EOS>
__global__ __launch_bounds__(256) void k_transport(const double *K, View Q, View st, View tr, View hsp, BoxList L,
                                                   int binary) {
  int i, j, k;
  if (!boxlist_cell(L, blockIdx.x, i, j, k)) return;
  const Gas<NS> g{K};
  transport_cell<NS, TRANS, EOS>(g, Q, st, Q.off(i, j, k), tr, hsp, tr.off(i, j, k), hsp.off(i, j, k), binary != 0);
}

// navstokes/kernels/dqdxyz.mako:16-41: the cell's gradients of u, v, w, T and the carried Y
template <int NS>
__device__ __forceinline__ void grads_cell(const View &Q, const View &st, const View &Jinv, const View &gr, int i, int j,
                                           int k) {
#pragma clang fp contract(off)
  constexpr int NGRAD = NS + 3;
  const int64_t o = Q.off(i, j, k);
  const int64_t d[3] = {1, Q.sj, Q.sk};
  double rp[3], rm[3], inv[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    rp[a] = Q.at(o + d[a], 0);
    rm[a] = Q.at(o - d[a], 0);
    inv[a] = 1.0 / (rp[a] * rm[a]);
  }
  const int64_t oj = Jinv.off(i, j, k), og = gr.off(i, j, k);
  double J[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int e = 0; e < 3; e++) J[a][e] = Jinv.at(oj, 3 * a + e);
  // every difference before the first store, so a thread waits on memory once
  double dd[NGRAD][3];
#pragma unroll
  for (int l = 0; l < NGRAD; l++) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (l == 3) {
        dd[l][a] = 0.5 * (st.at(o + d[a], 1) - st.at(o - d[a], 1));
      } else {
        const int m = l + 1;
        dd[l][a] = 0.5 * (Q.at(o + d[a], m) * rm[a] - Q.at(o - d[a], m) * rp[a]) * inv[a];
      }
    }
  }
#pragma unroll
  for (int l = 0; l < NGRAD; l++)
#pragma unroll
    for (int e = 0; e < 3; e++)
      gr.at(og, 3 * l + e) = dd[l][0] * J[0][e] + dd[l][1] * J[1][e] + dd[l][2] * J[2][e];
}

template <int NS>
__global__ __launch_bounds__(256) void k_grads(View Q, View st, View Jinv, View gr, Box3 b) {
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  grads_cell<NS>(Q, st, Jinv, gr, i, j, k);
}

// the transport over a box list and the gradients over a box in one launch: the transport's
// workgroups first, then the gradients' (flat, i fastest); neither reads what the other writes
template <int NS, int TRANS, int EOS>
__global__ __launch_bounds__(256) void k_transport_grads(const double *K, View Q, View st, View tr, View hsp,
                                                         BoxList L, unsigned tblocks, int binary, View Jinv, View gr,
                                                         Box3 gb) {
  int i, j, k;
  if (blockIdx.x < tblocks) {
    if (!boxlist_cell(L, blockIdx.x, i, j, k)) return;
    const Gas<NS> g{K};
    transport_cell<NS, TRANS, EOS>(g, Q, st, Q.off(i, j, k), tr, hsp, tr.off(i, j, k), hsp.off(i, j, k), binary != 0);
    return;
  }
  if (!flat_at(gb, blockIdx.x - tblocks, i, j, k)) return;
  grads_cell<NS>(Q, st, Jinv, gr, i, j, k);
}

struct DiffArgs {
  View Q, st, tr, hsp, gr, xcc, xf, S, F;
  double alpha;
  Box3 b;
};

// navstokes/kernels/flux/alpha-damping.mako:3-99, added to F
template <int NS, int AXIS>
__global__ __launch_bounds__(256) void k_diffflux(DiffArgs a) {
  int i, j, k;
  if (!flat_cell_of(a.b, i, j, k)) return;
  const int64_t off_ = a.F.off(i, j, k);
  double Fv[NS + 4];
#pragma unroll
  for (int l = 1; l < NS + 4; l++) Fv[l] = a.F.at(off_, l);
  Fv[0] = 0.0;
  diff_face<NS>(a.Q, a.st, a.tr, a.hsp, a.gr, a.xcc, a.xf, a.S, a.alpha, AXIS, i, j, k, Fv);
#pragma unroll
  for (int l = 1; l < NS + 4; l++) a.F.at(off_, l) = Fv[l];
}

}  // namespace pga

using namespace pga;

// the transport call's checks and its box list (each box on a workgroup's boundary); total the
// list's padded cell count
static int32_t transport_list(const char *who, const pga_case *c, pga_view Q, pga_view state, pga_view trstate,
                              pga_view hsp, const pga_box *boxes, int32_t nboxes, BoxList &L, int64_t &total) {
  if (!c || nboxes < 0 || nboxes > 8 || (nboxes && !boxes)) {
    set_error("%s: null argument or more than 8 boxes", who);
    return PGA_EINVAL;
  }
  const pga_case_desc &d = c->desc;
  if (d.trans < 0) {
    set_error("%s: the case has no transport model", who);
    return PGA_EUNSUP;
  }
  const int ne = d.ns + 4;
  for (int n = 0; n < nboxes; n++) {
    PGA_REQUIRE(box_ok(boxes[n]), "%s: box %d has a negative extent", who, n);
    PGA_REQUIRE(covers(Q, boxes[n], ne) && covers(state, boxes[n], nstate(d)) && covers(trstate, boxes[n], 2 + d.ns) &&
                    covers(hsp, boxes[n], d.ns),
                "%s: box %d reaches outside its arrays", who, n);
  }
  L = BoxList{};
  total = 0;
  for (int n = 0; n < nboxes; n++) {
    if (empty(boxes[n])) continue;
    L.b[L.n] = box3(boxes[n]);
    L.first[L.n] = total;
    total += ((int64_t)boxes[n].extent[0] * boxes[n].extent[1] * boxes[n].extent[2] + 255) / 256 * 256;
    L.n++;
  }
  PGA_REQUIRE(total < FLAT_MAX, "%s: %lld cells; at most 2^31", who, (long long)total);
  return PGA_OK;
}

extern "C" int32_t pga_transport(const pga_case *c, pga_view Q, pga_view state, pga_view trstate, pga_view hsp,
                                 const pga_box *boxes, int32_t nboxes, void *stream) {
  BoxList L;
  int64_t total;
  const int32_t rc = transport_list("pga_transport", c, Q, state, trstate, hsp, boxes, nboxes, L, total);
  if (rc != PGA_OK || !total) return rc;
  const pga_case_desc &d = c->desc;
  // every box in one launch
  const dim3 grid((unsigned)(total / 256));
  const bool done = dispatch_eos(d, [&]<int N, int E>() {
    if (d.trans == TRANS_CONSTANT)
      hipLaunchKernelGGL((k_transport<N, TRANS_CONSTANT, E>), grid, dim3(256), 0, stream_of(stream), c->K, view(Q),
                         view(state), view(trstate), view(hsp), L, d.diffusion);
    else if (d.trans == TRANS_KINETIC)
      hipLaunchKernelGGL((k_transport<N, TRANS_KINETIC, E>), grid, dim3(256), 0, stream_of(stream), c->K, view(Q),
                         view(state), view(trstate), view(hsp), L, d.diffusion);
    else if constexpr (E == EOS_REAL)
      hipLaunchKernelGGL((k_transport<N, TRANS_CHUNG, E>), grid, dim3(256), 0, stream_of(stream), c->K, view(Q),
                         view(state), view(trstate), view(hsp), L, d.diffusion);
  });
  if (!done) {
    set_error("pga_transport: no kernel for %d species with eos %d", d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_transport");
}

// the gradient call's checks
static int32_t grads_check(const char *who, const pga_case *c, pga_view Q, pga_view state, pga_view Jinv,
                           pga_view grads, pga_box box) {
  const pga_case_desc &d = c->desc;
  const int ne = d.ns + 4;
  const int one[3] = {1, 1, 1};
  PGA_REQUIRE(covers(Q, box, one, one, ne) && covers(state, box, one, one, nstate(d)),
              "%s: the stencil reaches outside Q or the state (one halo layer needed)", who);
  PGA_REQUIRE(covers(Jinv, box, 9) && covers(grads, box, 3 * (ne - 1)), "%s: the box reaches outside Jinv or grads", who);
  PGA_REQUIRE(Q.sj == state.sj && Q.sk == state.sk, "%s: Q and state must share their layout", who);
  return PGA_OK;
}

extern "C" int32_t pga_transport_grads(const pga_case *c, pga_view Q, pga_view state, pga_view trstate, pga_view hsp,
                                       const pga_box *boxes, int32_t nboxes, pga_view Jinv, pga_view grads,
                                       pga_box gbox, void *stream) {
  const char *who = "pga_transport_grads";
  if (!box_ok(gbox)) {
    set_error("%s: the gradients' box has a negative extent", who);
    return PGA_EINVAL;
  }
  BoxList L;
  int64_t total;
  int32_t rc = transport_list(who, c, Q, state, trstate, hsp, boxes, nboxes, L, total);
  if (rc != PGA_OK) return rc;
  const bool grad = !empty(gbox);
  if (grad && (rc = grads_check(who, c, Q, state, Jinv, grads, gbox)) != PGA_OK) return rc;
  const int64_t gcells = grad ? (int64_t)gbox.extent[0] * gbox.extent[1] * gbox.extent[2] : 0;
  PGA_REQUIRE(total + gcells < FLAT_MAX - FLAT, "%s: %lld cells in one launch; at most 2^31", who,
              (long long)(total + gcells));
  const unsigned tblocks = (unsigned)(total / 256), gblocks = (unsigned)((gcells + FLAT - 1) / FLAT);
  if (!tblocks && !gblocks) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const Box3 gb = grad ? box3(gbox) : Box3{};
  const dim3 grid(tblocks + gblocks);
  const bool done = dispatch_eos(d, [&]<int N, int E>() {
    if (d.trans == TRANS_CONSTANT)
      hipLaunchKernelGGL((k_transport_grads<N, TRANS_CONSTANT, E>), grid, dim3(256), 0, stream_of(stream), c->K,
                         view(Q), view(state), view(trstate), view(hsp), L, tblocks, d.diffusion, view(Jinv),
                         view(grads), gb);
    else if (d.trans == TRANS_KINETIC)
      hipLaunchKernelGGL((k_transport_grads<N, TRANS_KINETIC, E>), grid, dim3(256), 0, stream_of(stream), c->K,
                         view(Q), view(state), view(trstate), view(hsp), L, tblocks, d.diffusion, view(Jinv),
                         view(grads), gb);
    else if constexpr (E == EOS_REAL)
      hipLaunchKernelGGL((k_transport_grads<N, TRANS_CHUNG, E>), grid, dim3(256), 0, stream_of(stream), c->K,
                         view(Q), view(state), view(trstate), view(hsp), L, tblocks, d.diffusion, view(Jinv),
                         view(grads), gb);
  });
  if (!done) {
    set_error("%s: no kernel for %d species with eos %d", who, d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched(who);
}

extern "C" int32_t pga_grads(const pga_case *c, pga_view Q, pga_view state, pga_view Jinv, pga_view grads,
                             pga_box box, void *stream) {
  if (!c || !box_ok(box)) {
    set_error("pga_grads: null case or bad box");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int32_t rc = grads_check("pga_grads", c, Q, state, Jinv, grads, box);
  if (rc != PGA_OK) return rc;
  const Box3 b = box3(box);
  bool done = false;
#define X(N)                                                                                                    \
  if (d.ns == N) {                                                                                              \
    hipLaunchKernelGGL((k_grads<N>), grid_of(b), dim3(TI, TJ), 0, stream_of(stream), view(Q), view(state),      \
                       view(Jinv), view(grads), b);                                                             \
    done = true;                                                                                                \
  }
  PGA_NS_CASES(X)
#undef X
  if (!done) {
    set_error("pga_grads: no kernel for %d species", d.ns);
    return PGA_EUNSUP;
  }
  return launched("pga_grads");
}

extern "C" int32_t pga_diffflux(const pga_case *c, int32_t axis, double alpha, pga_view Q, pga_view state,
                                pga_view trstate, pga_view hsp, pga_view grads, pga_view xcc, pga_view xf, pga_view S,
                                pga_view F, pga_box box, void *stream) {
  if (!c || axis < 0 || axis > 2 || !box_ok(box)) {
    set_error("pga_diffflux: null case, an axis other than 0..2, or a bad box");
    return PGA_EINVAL;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int ne = d.ns + 4;
  int lo[3] = {0, 0, 0};
  const int hi[3] = {0, 0, 0};
  lo[axis] = 1;
  PGA_REQUIRE(covers(Q, box, lo, hi, ne) && covers(state, box, lo, hi, nstate(d)) && covers(trstate, box, lo, hi, 2 + d.ns) &&
                  covers(hsp, box, lo, hi, d.ns) && covers(grads, box, lo, hi, 3 * (ne - 1)) &&
                  covers(xcc, box, lo, hi, 3),
              "pga_diffflux: the face's cells reach outside a cell array");
  PGA_REQUIRE(covers(xf, box, 3) && covers(S, box, 3) && covers(F, box, ne), "pga_diffflux: the box reaches outside xf, S or F");
  PGA_REQUIRE(Q.sj == state.sj && Q.sk == state.sk && Q.sj == trstate.sj && Q.sk == trstate.sk && Q.sj == hsp.sj &&
                  Q.sk == hsp.sk && Q.sj == grads.sj && Q.sk == grads.sk && Q.sj == xcc.sj && Q.sk == xcc.sk,
              "pga_diffflux: the cell arrays must share their layout");
  if (!flat_fits(box)) return PGA_ERANGE;
  DiffArgs a{view(Q), view(state), view(trstate), view(hsp), view(grads), view(xcc), view(xf), view(S), view(F),
             alpha, box3(box)};
  bool done = false;
#define X(N)                                                                                                     \
  if (d.ns == N) {                                                                                               \
    flat_slabs(box3(box), [&](const Box3 &sb) {                                                                  \
      a.b = sb;                                                                                                  \
      if (axis == 0)                                                                                             \
        hipLaunchKernelGGL((k_diffflux<N, 0>), flat_grid(sb), dim3(FLAT), 0, stream_of(stream), a);             \
      else if (axis == 1)                                                                                        \
        hipLaunchKernelGGL((k_diffflux<N, 1>), flat_grid(sb), dim3(FLAT), 0, stream_of(stream), a);             \
      else                                                                                                       \
        hipLaunchKernelGGL((k_diffflux<N, 2>), flat_grid(sb), dim3(FLAT), 0, stream_of(stream), a);             \
    });                                                                                                          \
    done = true;                                                                                                 \
  }
  PGA_NS_CASES(X)
#undef X
  if (!done) {
    set_error("pga_diffflux: no kernel for %d species", d.ns);
    return PGA_EUNSUP;
  }
  return launched("pga_diffflux");
}

// plugins/simulator/kernels/viscsponge.mako: after the transport, the viscosity raised along a
// line from its origin to its end, to the multiplier times itself there and beyond
namespace pga {
struct SpongeLine {
  double nx, ny, nz, start, length, mult;
};

__global__ __launch_bounds__(256) void k_viscsponge(View xcc, View tr, SpongeLine s, Box3 b) {
#pragma clang fp contract(off)
  int i, j, k;
  if (!cell_of(b, i, j, k)) return;
  const int64_t ox = xcc.off(i, j, k);
  const double along = (xcc.at(ox, 0) * s.nx + xcc.at(ox, 1) * s.ny + xcc.at(ox, 2) * s.nz - s.start) / s.length;
  tr.at(tr.off(i, j, k), 0) *= 1.0 + fmin(fmax(0.0, along * (s.mult - 1.0)), s.mult - 1.0);
}
}  // namespace pga

extern "C" int32_t pga_viscsponge(pga_view xcc, pga_view trstate, const double *line, const pga_box *boxes,
                                  int32_t nboxes, void *stream) {
  if (!line || nboxes < 0 || nboxes > 8 || (nboxes && !boxes)) {
    set_error("pga_viscsponge: null argument or more than 8 boxes");
    return PGA_EINVAL;
  }
  if (!(line[4] > 0.0)) {
    set_error("pga_viscsponge: the sponge's length must be positive, not %g", line[4]);
    return PGA_EINVAL;
  }
  for (int n = 0; n < nboxes; n++)
    PGA_REQUIRE(box_ok(boxes[n]) && covers(xcc, boxes[n], 3) && covers(trstate, boxes[n], 1),
                "pga_viscsponge: box %d reaches outside xcc or trstate", n);
  const SpongeLine s{line[0], line[1], line[2], line[3], line[4], line[5]};
  for (int n = 0; n < nboxes; n++) {
    if (empty(boxes[n])) continue;
    const Box3 b = box3(boxes[n]);
    hipLaunchKernelGGL(k_viscsponge, grid_of(b), dim3(TI, TJ), 0, stream_of(stream), view(xcc), view(trstate), s, b);
  }
  return launched("pga_viscsponge");
}
