// The CFL reduction (euler/kernels/cflmax.mako:3-33, multiblock/solver.py:292-296): per interior
// cell the acoustic, convective and combined inverse time scales, folded by fmax from -inf
// (codegen/grammar.py:34-42). One flat launch writes a partial per workgroup into the caller's
// work array; one workgroup folds the partials into out[3]. fmax is exact and order-free, so the
// result does not depend on the launch shape.
#include <mutex>

#include "pga_common.hpp"
#include "pga_internal.hpp"

namespace pga {

struct CflArgs {
  View Q, st, d, S[3];
  Box3 b;
  double *partials;
};

// fmax of v over a 256-lane workgroup: down the 64 lanes of each wave, then across the 4 waves
__device__ __forceinline__ void block_fmax3(double (&v)[3], double (*lds)[3]) {
#pragma unroll
  for (int n = 0; n < 3; n++)
    for (int off = 32; off > 0; off >>= 1) v[n] = fmax(v[n], __shfl_down(v[n], off, 64));
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int n = 0; n < 3; n++) lds[wave][n] = v[n];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < FLAT / 64; w++)
      for (int n = 0; n < 3; n++) v[n] = fmax(v[n], lds[w][n]);
}

__global__ __launch_bounds__(FLAT) void k_cflmax(CflArgs a) {
#pragma clang fp contract(off)
  __shared__ double lds[FLAT / 64][3];
  double best[3] = {-__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
  int i, j, k;
  if (flat_cell_of(a.b, i, j, k)) {
    const int64_t o = a.Q.off(i, j, k), od = a.d.off(i, j, k), os = a.st.off(i, j, k);
    const double di = a.d.at(od, 0), dj = a.d.at(od, 1), dk = a.d.at(od, 2);
    const double rhoinv = 1.0 / a.Q.at(o, 0);
    const double u = a.Q.at(o, 1) * rhoinv, v = a.Q.at(o, 2) * rhoinv, w = a.Q.at(o, 3) * rhoinv;
    double un[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      const View &S = a.S[ax];
      const int64_t lo = S.off(i, j, k), hi = lo + (ax == 0 ? 1 : ax == 1 ? S.sj : S.sk);
      const double L0 = S.at(lo, 0), L1 = S.at(lo, 1), L2 = S.at(lo, 2);
      const double H0 = S.at(hi, 0), H1 = S.at(hi, 1), H2 = S.at(hi, 2);
      const double S0inv = 1.0 / sqrt(L0 * L0 + L1 * L1 + L2 * L2);
      const double S1inv = 1.0 / sqrt(H0 * H0 + H1 * H1 + H2 * H2);
      const double nx = 0.5 * (L0 * S0inv + H0 * S1inv);
      const double ny = 0.5 * (L1 * S0inv + H1 * S1inv);
      const double nz = 0.5 * (L2 * S0inv + H2 * S1inv);
      un[ax] = fabs(nx * u + ny * v + nz * w);
    }
    const double c = a.st.at(os, 7);
    // an infinite length contributes nothing
    const double acoustic = c * sqrt(1.0 / (di * di) + 1.0 / (dj * dj) + 1.0 / (dk * dk));
    const double convective = un[0] / di + un[1] / dj + un[2] / dk;
    best[0] = acoustic;
    best[1] = convective;
    best[2] = acoustic + convective;
  }
  block_fmax3(best, lds);
  if (threadIdx.x == 0)
    for (int n = 0; n < 3; n++) a.partials[3 * (int64_t)blockIdx.x + n] = best[n];
}

__global__ __launch_bounds__(FLAT) void k_cflfold(const double *__restrict__ partials, int64_t nparts,
                                                  double *__restrict__ out) {
  __shared__ double lds[FLAT / 64][3];
  double best[3] = {-__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
  for (int64_t p = threadIdx.x; p < nparts; p += FLAT)
    for (int n = 0; n < 3; n++) best[n] = fmax(best[n], partials[3 * p + n]);
  block_fmax3(best, lds);
  if (threadIdx.x == 0)
    for (int n = 0; n < 3; n++) out[n] = best[n];
}

// The allocation-free entry point's partials: one array per GPU for the process's life, addressed on
// the host only through the pointer hipGetSymbolAddress gives for the current device (the array's
// name in host code is a stand-in, not a device address). Every call on one GPU shares it.
constexpr int64_t CFL_SCRATCH_PARTS = 262144;  // 6 MiB: boxes up to 2^26 cells (a 406^3 block)
static __device__ __align__(256) double scratch[3 * CFL_SCRATCH_PARTS];

constexpr int CFL_MAX_DEVICES = 64;
static std::mutex scratch_lock;
static double *scratch_ptr[CFL_MAX_DEVICES];

// the current device's scratch array, its address looked up once per device; null with the error set
static double *device_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= CFL_MAX_DEVICES) {
    set_error("pga_cflmax_noalloc: no current device below %d", CFL_MAX_DEVICES);
    return nullptr;
  }
  std::lock_guard<std::mutex> hold(scratch_lock);
  if (!scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("pga_cflmax_noalloc: the scratch array's address on device %d: %s", dev, hipGetErrorString(e));
      return nullptr;
    }
    scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return scratch_ptr[dev];
}

// the partials one call writes: a workgroup per FLAT cells of each slab
static int64_t cfl_parts(const pga_box &box) {
  int64_t parts = 0;
  if (!empty(box)) flat_slabs(box3(box), [&](const Box3 &sb) { parts += flat_grid(sb).x; });
  return parts;
}

}  // namespace pga

using namespace pga;

extern "C" int64_t pga_cflmax_worksize(pga_box box) { return box_ok(box) ? 3 * std::max<int64_t>(cfl_parts(box), 1) : -1; }

extern "C" int32_t pga_cflmax(const pga_case *c, pga_view Q, pga_view state, pga_view dijk, pga_view iS, pga_view jS,
                              pga_view kS, pga_box box, double *work, int64_t nwork, double *out, void *stream) {
  if (!c || !box_ok(box) || !work || !out) {
    set_error("pga_cflmax: null case, work or out, or a bad box");
    return PGA_EINVAL;
  }
  if (!flat_fits(box)) return PGA_ERANGE;
  const int64_t parts = cfl_parts(box);
  PGA_REQUIRE(nwork >= 3 * std::max<int64_t>(parts, 1), "pga_cflmax: work holds %lld doubles, the box needs %lld",
              (long long)nwork, (long long)(3 * std::max<int64_t>(parts, 1)));
  const int ne = c->desc.ns + 4;
  if (!empty(box)) {
    int lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    PGA_REQUIRE(covers(Q, box, ne) && covers(state, box, nstate(c->desc)) && covers(dijk, box, 3),
                "pga_cflmax: the box reaches outside Q, state or dijk");
    const pga_view *S[3] = {&iS, &jS, &kS};
    for (int a = 0; a < 3; a++) {
      int h[3] = {0, 0, 0};
      h[a] = hi[a];
      PGA_REQUIRE(covers(*S[a], box, lo, h, 3), "pga_cflmax: the cells' faces reach outside iS, jS or kS");
    }
  }
  CflArgs a{view(Q), view(state), view(dijk), {view(iS), view(jS), view(kS)}, box3(box), work};
  int64_t at = 0;
  if (!empty(box))
    flat_slabs(box3(box), [&](const Box3 &sb) {
      a.b = sb;
      a.partials = work + 3 * at;
      const dim3 grid = flat_grid(sb);
      hipLaunchKernelGGL(k_cflmax, grid, dim3(FLAT), 0, stream_of(stream), a);
      at += grid.x;
    });
  hipLaunchKernelGGL(k_cflfold, dim3(1), dim3(FLAT), 0, stream_of(stream), work, parts, out);
  return launched("pga_cflmax");
}

extern "C" int32_t pga_cflmax_noalloc(const pga_case *c, pga_view Q, pga_view state, pga_view dijk, pga_view iS,
                                      pga_view jS, pga_view kS, pga_box box, double *out, void *stream) {
  if (!c || !box_ok(box) || !out) {
    set_error("pga_cflmax_noalloc: null case or out, or a bad box");
    return PGA_EINVAL;
  }
  if (!flat_fits(box)) return PGA_ERANGE;
  const int64_t parts = cfl_parts(box);
  PGA_REQUIRE(parts <= CFL_SCRATCH_PARTS,
              "pga_cflmax_noalloc: the box needs %lld partials, the scratch array holds %lld (boxes up to %lld cells)",
              (long long)parts, (long long)CFL_SCRATCH_PARTS, (long long)(CFL_SCRATCH_PARTS * FLAT));
  double *work = device_scratch();
  if (!work) return PGA_EHIP;
  return pga_cflmax(c, Q, state, dijk, iS, jS, kS, box, work, 3 * CFL_SCRATCH_PARTS, out, stream);
}
