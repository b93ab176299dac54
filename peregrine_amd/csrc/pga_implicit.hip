// Implicit chemistry's entry points (the kernels are in pga_implicit.hpp): per part of the step,
// the forward-difference matrices, their batched FP32 LU on the matrix cores, and the refined
// linearly implicit step; and the batched solve of caller-given systems.
#include "pga_implicit.hpp"
#include "pga_internal.hpp"

namespace pga {

// the batched solve's own kernels: A (row-major FP64) into the factor array as FP32, and the
// refinement of each system against A
__global__ __launch_bounds__(256) void k_pack(const double *__restrict__ A, int nv, int64_t n, float *__restrict__ W) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n * IM_E) return;
  const int64_t sys = t / IM_E;
  const int e = (int)(t % IM_E), r = e / IM_N, c = e % IM_N;
  W[lu_at(sys, r, c)] = (r < nv && c < nv) ? (float)A[(sys * nv + r) * nv + c] : (r == c ? 1.0f : 0.0f);
}

template <int NV>
__global__ __launch_bounds__(256) void k_refine(const double *__restrict__ A, const double *__restrict__ bv,
                                                double *__restrict__ x, int32_t *__restrict__ sweeps, int64_t n,
                                                int maxsweeps, double tol, const float *__restrict__ W) {
  const int64_t sys = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (sys >= n) return;
  double b[NV], xs[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) b[m] = bv[sys * NV + m];
  const int taken = refine<NV>(W, A, [](int64_t c, int r, int m) { return (c * NV + r) * NV + m; }, sys, b, xs,
                               maxsweeps, tol);
#pragma unroll
  for (int m = 0; m < NV; m++) x[sys * NV + m] = xs[m];
  if (sweeps) sweeps[sys] = taken;
}

static int64_t im_cells(const pga_box &b) { return (int64_t)b.extent[0] * b.extent[1] * b.extent[2]; }

// doubles of a cell's cache of f at one T for the case's mechanism (ImCache), -1 for none
int64_t im_cache_doubles(const pga_case_desc &d) {
  int64_t n = -1;
#define X(NAME, IDX) \
  if (d.mechanism == IDX) n = im_cache_n<Mech_##NAME>;
  PGA_MECHANISMS(X)
#undef X
  return n;
}
static int64_t im_groups(int64_t cells) { return (cells + IM_G - 1) / IM_G; }

template <int NB>
static void launch_lu(float *W, int64_t n, hipStream_t s) {
  const int64_t blocks = (im_groups(n) + 3) / 4;
  hipLaunchKernelGGL((k_lu<NB>), dim3((unsigned)blocks), dim3(256), 0, s, W, n);
}

}  // namespace pga

using namespace pga;

extern "C" int64_t pga_implicit_worksize(const pga_case *c, pga_box box) {
  if (!c || !box_ok(box) || c->desc.ns + 1 > IM_N) return -1;
  const int64_t cells = im_cells(box), groups = im_groups(cells), nv = c->desc.ns + 1;
  const int64_t nt = im_cache_doubles(c->desc);
  if (nt < 0) return -1;
  return groups * nv * nv * IM_G + (3 * nv + nt) * cells;
}

extern "C" int32_t pga_implicit(const pga_case *c, const pga_implicit_opts *o, pga_view Q, pga_view state,
                                pga_view out, int32_t begin_dQ, int32_t *sweeps, int64_t nsweeps, double *work,
                                int64_t nwork, const double *dt, pga_box box, void *stream) {
  if (!c || !o || !box_ok(box) || !dt || !sweeps || !work) {
    set_error("pga_implicit: null case, options, dt, sweeps or work, or a bad box");
    return PGA_EINVAL;
  }
  if (o->parts < 1 || o->max_sweeps < 1 || !(o->tolerance > 0.0)) {
    set_error("pga_implicit: parts >= 1, max_sweeps >= 1 and a tolerance above zero are required");
    return PGA_EINVAL;
  }
  if (c->desc.mechanism < 0 || !c->chem) {
    set_error("pga_implicit: the case has no compiled mechanism");
    return PGA_EUNSUP;
  }
  if (c->desc.ns + 1 > IM_N) {
    set_error("pga_implicit: %d species and T exceed the %d x %d matrices of the batched LU", c->desc.ns, IM_N, IM_N);
    return PGA_EUNSUP;
  }
  if (empty(box)) return PGA_OK;
  const pga_case_desc &d = c->desc;
  const int64_t cells = im_cells(box), groups = im_groups(cells), nv = d.ns + 1;
  PGA_REQUIRE(cells < FLAT_MAX - FLAT, "pga_implicit: a box of %lld cells; at most 2^31", (long long)cells);
  PGA_REQUIRE(covers(Q, box, d.ns + 4) && covers(state, box, nstate(d)) &&
                  covers(out, box, begin_dQ ? d.ns + 4 : d.ns - 1),
              "pga_implicit: the box reaches outside Q, state or out");
  PGA_REQUIRE(nsweeps >= cells, "pga_implicit: sweeps holds %lld counts, the box has %lld cells", (long long)nsweeps,
              (long long)cells);
  const int64_t need = pga_implicit_worksize(c, box);
  PGA_REQUIRE(nwork >= need, "pga_implicit: work holds %lld doubles, the box needs %lld", (long long)nwork,
              (long long)need);
  ImArgs a{};
  a.K = c->K, a.C = c->chem;
  a.Q = view(Q), a.st = view(state), a.out = view(out);
  a.begin_dQ = begin_dQ != 0;
  a.A = work;
  a.y = a.A + groups * nv * nv * IM_G;
  a.tc = a.y + 3 * nv * cells;
  a.ncell = cells, a.sweeps = sweeps, a.dt = dt, a.b = box3(box);
  a.parts = o->parts, a.maxsweeps = o->max_sweeps, a.tol = o->tolerance;
  const dim3 grid = flat_grid(a.b);
  hipStream_t s = stream_of(stream);
  bool done = false;
#define X(NAME, IDX)                                                                              \
  if (d.mechanism == IDX) {                                                                       \
    constexpr int NV = Mech_##NAME::NS + 1, NB = (NV + 3) / 4;                                    \
    for (int part = 0; part < o->parts; part++) {                                                 \
      a.first = part == 0, a.last = part == o->parts - 1;                                         \
      hipLaunchKernelGGL((k_itc<Mech_##NAME>), grid, dim3(FLAT), 0, s, a);                       \
      hipLaunchKernelGGL((k_ibase<Mech_##NAME>), grid, dim3(FLAT), 0, s, a);                     \
      hipLaunchKernelGGL((k_icol<Mech_##NAME>), dim3(grid.x * (NV + 1)), dim3(FLAT), 0, s, a);   \
      hipLaunchKernelGGL((k_lusolve<NB, NV>), dim3((unsigned)((groups + 3) / 4)), dim3(256), 0, s, a); \
    }                                                                                             \
    hipLaunchKernelGGL((k_iend<Mech_##NAME>), grid, dim3(FLAT), 0, s, a);                        \
    done = true;                                                                                  \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_implicit: no mechanism %d", d.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_implicit");
}

extern "C" int64_t pga_lu_worksize(int64_t n, int32_t nv) {
  if (n < 0 || nv < 1 || nv > IM_N) return -1;
  return im_groups(n) * (IM_GROUP / 2);
}

extern "C" int32_t pga_lu_solve(int64_t n, int32_t nv, const double *A, const double *b, double *x,
                                int32_t max_sweeps, double tolerance, int32_t *sweeps, double *work, int64_t nwork,
                                void *stream) {
  if (n < 0 || nv < 1 || nv > IM_N || max_sweeps < 1 || !(tolerance >= 0.0) ||
      (n && (!A || !b || !x || !work))) {
    set_error("pga_lu_solve: n >= 0 systems of 1..%d unknowns, max_sweeps >= 1, a tolerance >= 0, and A, b, x and "
              "work", IM_N);
    return PGA_EINVAL;
  }
  if (!n) return PGA_OK;
  PGA_REQUIRE(n * IM_E < FLAT_MAX, "pga_lu_solve: %lld systems; at most 2^31 entries", (long long)n);
  PGA_REQUIRE(nwork >= pga_lu_worksize(n, nv), "pga_lu_solve: work holds %lld doubles, %lld systems need %lld",
              (long long)nwork, (long long)n, (long long)pga_lu_worksize(n, nv));
  float *W = reinterpret_cast<float *>(work);
  hipStream_t s = stream_of(stream);
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((n * IM_E + 255) / 256)), dim3(256), 0, s, A, nv, n, W);
  const dim3 grid((unsigned)((n + 255) / 256));
  sfor<1, IM_N + 1>([&]<int NV>() {
    if (nv != NV) return;
    launch_lu<(NV + 3) / 4>(W, n, s);
    hipLaunchKernelGGL((k_refine<NV>), grid, dim3(256), 0, s, A, b, x, sweeps, n, (int)max_sweeps, tolerance,
                       (const float *)W);
  });
  return launched("pga_lu_solve");
}
