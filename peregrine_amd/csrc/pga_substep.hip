// Sub-stepped chemistry (chemint/substepped.py:11-68, chemint/kernels/{substep,bulk,stiff,reset}.mako):
// each cell integrates its own chemistry over the step in explicit substeps, each as long as the
// limiting species allows, and begins dQ with the mean rate over the step.
//
// Two passes when the bulk limit is below the cap: the bulk pass runs every cell up to `bulk`
// substeps; a cell that is not done is listed, in one of `bins` bins by its substep count at the
// previous right-hand side, and the stiff pass runs the listed cells from the start up to `cap`,
// bin after bin, so the lanes of a wave take cells of like cost. The stiff pass restarts rather
// than resumes, and without the equilibrium exit (only a cell's first substep judges it, and a
// listed cell took that substep in the bulk pass), so every cell's result is the one-pass form's
// to the bit. The list is filled with vector atomics; its order varies, the results do not.
#include "generated/mechanisms.hpp"
#include "pga_chem.hpp"
#include "pga_internal.hpp"

namespace pga {

constexpr double SUB_STEPFRAC = 1e-12;  // a substep below this share of the step is none
constexpr int SUB_MAXBINS = 16;

struct SubArgs {
  const double *K, *C;
  View Q, st, out;
  int begin_dQ;
  int32_t *substeps, *list;
  int64_t seg;  // a bin's segment: as many slots as the box has cells
  const double *dt;
  Box3 b;
  int cap, bulk, bins, compacted;
  double eqtol, limitinv, binscale;
};

// substep.mako:5-83: up to `limit` substeps from (Y, T); src gathers the mean rate over the step.
// Returns the substeps taken; `remaining` is the share of the step left
template <class M, bool EXIT>
__device__ int integrate(const Gas<M::NS> &g, const double *__restrict__ C, int limit, double eqtol, double rho,
                         double rhoinv, double step, double (&Y)[M::NS], double &T, double (&src)[M::NS],
                         double &remaining) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS;
  const double eps = 2.220446049250313e-16, tiny = 2.2250738585072014e-308, stepinv = 1.0 / step;
  double rate[NS], D[NS];
  int sub = 0;
  bool early = false;
#pragma unroll 1
  for (; sub < limit && remaining > SUB_STEPFRAC; sub++) {
    typename M::Rates s;
    M::ratestate(g, C, rho, Y, T, s);
    // the first substep's fold carries the destruction rates the exit relaxes by
    if (EXIT && sub == 0)
      M::template fold<true>(C, s, rate, D);
    else
      M::template fold<false>(C, s, rate, D);
    // the limiting species by cross-multiplying, one quotient at the end; a trace species counts
    double abest = 0.0, bbest = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) {
      const double sign = copysign(1.0, rate[n]);
      const double headroom = 0.5 * ((1.0 - sign) * Y[n] + (1.0 + sign) * (1.0 - Y[n]));
      const double a = rho * fmax(headroom, eps), b = fmax(fabs(rate[n]), tiny);
      if (n == 0 || a * bbest < abest * b) {
        abest = a;
        bbest = b;
      }
    }
    const double share = fmin(remaining, abest / bbest * stepinv);
    // the temperature's rate at constant volume: cv, and the internal energy each species' rate carries
    double cv = 0.0, dTdt = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) {
      const double cpR = cpR_species(g, n, s.logT);
      const double e = e_species(g, n, T, s.logT, s.Tinv);
      cv += (cpR - 1.0) * g.Rsp(n) * Y[n];
      dTdt -= e * rate[n];
    }
    const double cvrhoinv = 1.0 / (cv * rho);
    dTdt *= cvrhoinv;
    if constexpr (EXIT) {
      // the equilibrium exit, judged at the first substep only: nothing moves past tolerance over
      // what remains, so the rest is one move with each species relaxed at its destruction
      // frequency
      const double tleft = remaining * step, tleftinv = 1.0 / tleft;
      bool settled = sub == 0 && tleft > 0.0 && fabs(dTdt) * tleft < 0.1;
#pragma unroll
      for (int n = 0; n < NS; n++) settled = settled && fabs(rate[n] * rhoinv) * tleft < eqtol;
      if (settled) {
        const double cfloor = C[M::CFLOOR];
        double dT = 0.0;
#pragma unroll
        for (int n = 0; n < NS; n++) {
          const double rhoden = rho * fmax(Y[n], cfloor * g.MW(n) * rhoinv);
          const double move = fmax((rate[n] * rhoinv) * tleft * rhoden / (rhoden + D[n] * tleft), -Y[n]);
          const double r = move * rho * tleftinv;
          const double e = e_species(g, n, T, s.logT, s.Tinv);
          dT -= e * r;
          Y[n] += move;
          src[n] += r * remaining;
        }
        T += dT * cvrhoinv * tleft;
        remaining = 0.0;
        early = true;
        break;
      }
    }
    const double tsub = share * step;
#pragma unroll
    for (int n = 0; n < NS; n++) {
      Y[n] += (rate[n] * rhoinv) * tsub;
      src[n] += rate[n] * share;
    }
    T += dTdt * tsub;
    remaining -= share;
    if (share <= SUB_STEPFRAC) {
      remaining = 0.0;
      early = true;
      break;
    }
  }
  return early ? sub + 1 : sub;
}

// one cell from its start (substep.mako:85-93 startstate) to its source (cell.mako:14-19
// beginsource); false when it is left for the stiff pass
template <class M, bool EXIT>
__device__ __forceinline__ bool cell_substeps(const SubArgs &a, int i, int j, int k, int limit, double &remaining,
                                              int &taken, double (&src)[M::NS]) {
  constexpr int NS = M::NS;
  const Gas<NS> g{a.K};
  const int64_t o = a.Q.off(i, j, k);
  const double rho = a.Q.at(o, 0), rhoinv = 1.0 / rho, step = *a.dt;
  double Y[NS];
  massfrac<NS>(a.Q, o, rhoinv, Y);
#pragma unroll
  for (int n = 0; n < NS; n++) src[n] = 0.0;
  double T = a.st.at(a.st.off(i, j, k), 1);
  remaining = 1.0;
  taken = integrate<M, EXIT>(g, a.C, limit, a.eqtol, rho, rhoinv, step, Y, T, src, remaining);
  return true;
}

template <int NS>
__device__ __forceinline__ void begin_source(const SubArgs &a, int i, int j, int k, const double (&src)[NS]) {
  const int64_t oo = a.out.off(i, j, k);
  if (a.begin_dQ) {
#pragma unroll
    for (int l = 0; l < 5; l++) a.out.at(oo, l) = 0.0;
#pragma unroll
    for (int n = 0; n < NS - 1; n++) a.out.at(oo, 5 + n) = src[n];
  } else {
#pragma unroll
    for (int n = 0; n < NS - 1; n++) a.out.at(oo, n) = src[n];
  }
}

__device__ __forceinline__ int64_t box_code(const Box3 &b, int i, int j, int k) {
  return (i - b.s0) + (int64_t)b.e0 * ((j - b.s1) + (int64_t)b.e1 * (k - b.s2));
}

// the list's header: [0] every listed cell, [1 + b] bin b's count (reset.mako)
__global__ void k_substep_reset(int32_t *list, int bins) {
  if ((int)threadIdx.x <= bins) list[threadIdx.x] = 0;
}

// bulk.mako:3-34
template <class M, bool EXIT>
__global__ __launch_bounds__(FLAT) void k_substep_bulk(SubArgs a) {
  constexpr int NS = M::NS;
  int i, j, k;
  if (!flat_cell_of(a.b, i, j, k)) return;
  double src[NS], remaining;
  int taken;
  cell_substeps<M, EXIT>(a, i, j, k, a.bulk, remaining, taken, src);
  const int64_t code = box_code(a.b, i, j, k);
  if (a.compacted && remaining > SUB_STEPFRAC) {
    int b = 0;
    if (a.bins > 1) {
      // by the cell's count from the last right-hand side, a cell just turned stiff to the top bin
      const int prev = a.substeps[code];
      const int expected = prev > a.bulk ? prev : a.cap;
      const int spaced = (int)(log((double)expected * a.limitinv) * a.binscale);
      b = spaced < 0 ? 0 : (spaced >= a.bins ? a.bins - 1 : spaced);
    }
    const int slot = atomicAdd(&a.list[1 + b], 1);
    atomicAdd(&a.list[0], 1);
    a.list[1 + a.bins + b * a.seg + slot] = (int32_t)code;
    return;
  }
  a.substeps[code] = taken;
  begin_source<NS>(a, i, j, k, src);
}

// stiff.mako:3-18: the listed cells, bin after bin, each from the start up to the cap. The grid
// covers every cell of the box; the lanes past the count the bulk pass left in the header return at
// once (a grid-stride loop around the integration would hold its state in scratch)
template <class M>
__global__ __launch_bounds__(FLAT) void k_substep_stiff(SubArgs a) {
  constexpr int NS = M::NS;
  const int total = a.list[0];
  {
    const int t = blockIdx.x * FLAT + threadIdx.x;
    if (t >= total) return;
    int cum = 0, entry = -1;
    for (int b = 0; b < a.bins; b++) {
      const int cnt = a.list[1 + b];
      if (t < cum + cnt) {
        entry = a.list[1 + a.bins + b * a.seg + (t - cum)];
        break;
      }
      cum += cnt;
    }
    if (entry < 0) return;
    const int i = a.b.s0 + (int)(entry % a.b.e0);
    const int j = a.b.s1 + (int)((entry / a.b.e0) % a.b.e1);
    const int k = a.b.s2 + (int)(entry / ((int64_t)a.b.e0 * a.b.e1));
    double src[NS], remaining;
    int taken;
    cell_substeps<M, false>(a, i, j, k, a.cap, remaining, taken, src);
    a.substeps[entry] = taken;
    begin_source<NS>(a, i, j, k, src);
  }
}

static int64_t box_cells(const pga_box &b) { return (int64_t)b.extent[0] * b.extent[1] * b.extent[2]; }

static bool opts_ok(const pga_substep_opts *o) {
  if (!o) return false;
  const int cap = o->max_substeps, bulk = o->bulk_substeps, bins = o->stiff_bins;
  const double tol = o->equilibrium_tolerance;
  return cap >= 1 && bulk >= 1 && bulk <= cap && bins >= 1 && bins <= SUB_MAXBINS && tol >= 0.0 && tol < 1.0;
}

}  // namespace pga

using namespace pga;

extern "C" int64_t pga_substep_listsize(const pga_substep_opts *o, pga_box box) {
  if (!opts_ok(o) || !box_ok(box)) return -1;
  return 1 + o->stiff_bins + (int64_t)o->stiff_bins * box_cells(box);
}

extern "C" int32_t pga_substep(const pga_case *c, const pga_substep_opts *o, pga_view Q, pga_view state, pga_view out,
                               int32_t begin_dQ, int32_t *substeps, int64_t nsubsteps, int32_t *list, int64_t nlist,
                               const double *dt, pga_box box, void *stream) {
  if (!c || !box_ok(box) || !dt || !substeps) {
    set_error("pga_substep: null case, dt or substeps, or a bad box");
    return PGA_EINVAL;
  }
  if (!opts_ok(o)) {
    set_error("pga_substep: max-substeps >= 1, 1 <= bulk-substeps <= max-substeps, 1 <= stiff-bins <= %d and "
              "0 <= equilibrium-tolerance < 1 are required", SUB_MAXBINS);
    return PGA_EINVAL;
  }
  if (c->desc.mechanism < 0 || !c->chem) {
    set_error("pga_substep: the case has no compiled mechanism");
    return PGA_EUNSUP;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int64_t cells = box_cells(box);
  PGA_REQUIRE(cells < FLAT_MAX - FLAT, "pga_substep: a box of %lld cells; at most 2^31 are listed",
              (long long)cells);
  PGA_REQUIRE(covers(Q, box, d.ns + 4) && covers(state, box, nstate(d)) &&
                  covers(out, box, begin_dQ ? d.ns + 4 : d.ns - 1),
              "pga_substep: the box reaches outside Q, state or out");
  PGA_REQUIRE(nsubsteps >= cells, "pga_substep: substeps holds %lld counts, the box has %lld cells",
              (long long)nsubsteps, (long long)cells);
  const bool compacted = o->bulk_substeps < o->max_substeps;
  if (compacted)
    PGA_REQUIRE(list && nlist >= pga_substep_listsize(o, box),
                "pga_substep: the stiff list holds %lld entries, the box needs %lld", (long long)nlist,
                (long long)pga_substep_listsize(o, box));
  SubArgs a{};
  a.K = c->K, a.C = c->chem;
  a.Q = view(Q), a.st = view(state), a.out = view(out);
  a.begin_dQ = begin_dQ != 0;
  a.substeps = substeps, a.list = list, a.seg = cells, a.dt = dt, a.b = box3(box);
  a.cap = o->max_substeps, a.bulk = o->bulk_substeps, a.bins = o->stiff_bins, a.compacted = compacted;
  a.eqtol = o->equilibrium_tolerance;
  // substepped.py:44-49: the bins' spacing in ln(count / limit)
  a.limitinv = 1.0 / a.bulk;
  a.binscale = a.bulk < a.cap ? a.bins / std::log((double)a.cap / a.bulk) : 0.0;
  const bool exit = a.eqtol > 0.0;
  const dim3 grid = flat_grid(a.b);
  hipStream_t s = stream_of(stream);
  bool done = false;
#define X(NAME, IDX)                                                                                           \
  if (d.mechanism == IDX) {                                                                                    \
    if (compacted) hipLaunchKernelGGL(k_substep_reset, dim3(1), dim3(64), 0, s, list, a.bins);                \
    if (exit)                                                                                                  \
      hipLaunchKernelGGL((k_substep_bulk<Mech_##NAME, true>), grid, dim3(FLAT), 0, s, a);                     \
    else                                                                                                       \
      hipLaunchKernelGGL((k_substep_bulk<Mech_##NAME, false>), grid, dim3(FLAT), 0, s, a);                    \
    if (compacted) hipLaunchKernelGGL((k_substep_stiff<Mech_##NAME>), grid, dim3(FLAT), 0, s, a);             \
    done = true;                                                                                               \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_substep: no mechanism %d", d.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_substep");
}
