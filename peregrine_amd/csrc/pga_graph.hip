// Graph capture and replay of a stream's launches, and the all-finite reduction, in the C ABI: what a
// binding of the reference needs to replay a step and to run its nancheck plugin without torch.
//
// The reference captures a host-free stretch of launches on its kernels stream with hipStreamBeginCapture
// / hipStreamEndCapture / hipGraphInstantiate and replays it with hipGraphLaunch
// (peregrine/backend/device/graph.py:10-47, peregrine/backend/hip/driver.py:159-176). The library's own
// launches record no memset, allocation or host synchronisation, so a step's launches capture whole.
//
// allfinite (plugins/solver/kernels/allfinite.mako:3-19): out[0] is set to 1 by one thread, then any
// thread whose cell holds a non-finite component stores 0 (every such store writes the same value);
// no memset, so the pair is graph-capturable too.
#include "pga_internal.hpp"

struct pga_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace pga {

__global__ void k_finite_init(int32_t *out) {
  if (threadIdx.x == 0) out[0] = 1;
}

__global__ __launch_bounds__(FLAT) void k_allfinite(View A, int ncomp, Box3 b, int32_t *out) {
  int i, j, k;
  if (!flat_cell_of(b, i, j, k)) return;
  const int64_t o = A.off(i, j, k);
  bool ok = true;
  for (int c = 0; c < ncomp; c++) ok = ok && isfinite(A.at(o, c));
  if (!ok) out[0] = 0;
}

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_graph_begin(void *stream) {
  if (!stream) {
    set_error("pga_graph_begin: capture needs a stream of its own (the null stream cannot be captured)");
    return PGA_EINVAL;
  }
  const hipError_t e = hipStreamBeginCapture(stream_of(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    set_error("pga_graph_begin: %s", hipGetErrorString(e));
    return PGA_EHIP;
  }
  return PGA_OK;
}

extern "C" int32_t pga_graph_end(void *stream, pga_graph **out) {
  if (!stream || !out) {
    set_error("pga_graph_end: a stream and an out pointer");
    return PGA_EINVAL;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(stream_of(stream), &g);
  if (e != hipSuccess || !g) {
    set_error("pga_graph_end: %s", hipGetErrorString(e));
    return PGA_EHIP;
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    set_error("pga_graph_end: instantiate: %s", hipGetErrorString(e));
    return PGA_EHIP;
  }
  *out = new pga_graph{g, x};
  return PGA_OK;
}

extern "C" int32_t pga_graph_launch(const pga_graph *g, void *stream) {
  if (!g) {
    set_error("pga_graph_launch: no graph");
    return PGA_EINVAL;
  }
  const hipError_t e = hipGraphLaunch(g->exec, stream_of(stream));
  if (e != hipSuccess) {
    set_error("pga_graph_launch: %s", hipGetErrorString(e));
    return PGA_EHIP;
  }
  return PGA_OK;
}

extern "C" int32_t pga_graph_destroy(pga_graph *g) {
  if (!g) return PGA_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return PGA_OK;
}

extern "C" int32_t pga_allfinite(pga_view A, int32_t ncomp, pga_box box, int32_t *out, void *stream) {
  if (!out || ncomp < 1 || !box_ok(box)) {
    set_error("pga_allfinite: an out pointer, one or more components and a box");
    return PGA_EINVAL;
  }
  hipLaunchKernelGGL(k_finite_init, dim3(1), dim3(64), 0, stream_of(stream), out);
  if (!empty(box)) {
    PGA_REQUIRE(covers(A, box, ncomp), "pga_allfinite: the box reaches outside the array");
    if (!flat_fits(box)) return PGA_ERANGE;
    const Box3 b = box3(box);
    flat_slabs(b, [&](const Box3 &sb) {
      hipLaunchKernelGGL(k_allfinite, flat_grid(sb), dim3(FLAT), 0, stream_of(stream), view(A), (int)ncomp, sb, out);
    });
  }
  return launched("pga_allfinite");
}
