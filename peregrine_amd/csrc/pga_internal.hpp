// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/peregrine_amd.h"
#include "pga_common.hpp"

struct pga_case {
  pga_case_desc desc;
  const double *K;     // the species constant table (device)
  const double *chem;  // the mechanism's constant table (device), or null
};

namespace pga {

void set_error(const char *fmt, ...);

inline hipStream_t stream_of(void *s) { return reinterpret_cast<hipStream_t>(s); }

// the status of the launch just made on this thread
inline int32_t launched(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return PGA_EHIP;
  }
  return PGA_OK;
}

// the elements of one component plane a view may hold: 4 GiB of doubles
constexpr int64_t PLANE_MAX = int64_t(1) << 29;

inline bool box_ok(const pga_box &b) {
  for (int a = 0; a < 3; a++)
    if (b.start[a] < 0 || b.extent[a] < 0) return false;
  return b.extent[2] <= 65535;
}

// a flat launch takes a box whose every k-plane indexes in 32 bits
inline bool flat_fits(const pga_box &b) {
  if ((int64_t)b.extent[0] * b.extent[1] < FLAT_MAX - FLAT) return true;
  set_error("a plane of %d x %d cells exceeds the flat launch's 2^31 cells", b.extent[0], b.extent[1]);
  return false;
}

inline bool empty(const pga_box &b) { return b.extent[0] == 0 || b.extent[1] == 0 || b.extent[2] == 0; }

// whether the cells [start - lo, start + extent + hi) along each axis lie inside the view,
// and it holds at least ncomp components
inline bool covers(const pga_view &v, const pga_box &b, const int lo[3], const int hi[3], int ncomp) {
  if (!v.data || v.n[3] < ncomp) return false;
  for (int a = 0; a < 3; a++) {
    if (b.extent[a] == 0) continue;
    if (b.start[a] - lo[a] < 0 || b.start[a] + b.extent[a] + hi[a] > v.n[a]) return false;
  }
  // the strides must not let one plane overlap the next, and a plane is addressed by 32-bit byte
  // offsets from its base (PLANE_MAX elements)
  return v.sj >= v.n[0] && v.sk >= v.sj * v.n[1] && v.sc >= v.sk * v.n[2] && v.sk * v.n[2] <= PLANE_MAX;
}

inline bool covers(const pga_view &v, const pga_box &b, int ncomp) {
  const int z[3] = {0, 0, 0};
  return covers(v, b, z, z, ncomp);
}

// Where a placed box reads its source (exchange/collection.py:55-96, the neighbour's window with
// its axes remapped): the destination cell start + (a0, a1, a2) takes the source cell
// s + sum_m a_m * (+-1 along source axis A_m), where orient[m] (1..6) names A_m (1, 2, 3 for
// +i, +j, +k; 4, 5, 6 for -i, -j, -k), the interface orientation digits of the reference; null
// is 1 2 3. As element offsets: base, the source cell s, and the step per destination axis.
struct Placement {
  int64_t base, step[3];
};

// false (with the error set) when orient is not three distinct axes or the source cells the box
// maps to reach outside src
inline bool placement(const char *who, const pga_view &src, const int32_t *s, const int32_t *orient,
                      const int32_t *ext, int ncomp, Placement &p) {
  int lo[3] = {0, 0, 0}, n[3] = {0, 0, 0};
  bool seen[3] = {false, false, false};
  const int64_t stride[3] = {1, src.sj, src.sk};
  for (int m = 0; m < 3; m++) {
    const int code = orient ? orient[m] : m + 1;
    if (code < 1 || code > 6 || seen[(code - 1) % 3]) {
      set_error("%s: orientation digits %d %d %d do not name three distinct axes", who, orient ? orient[0] : 1,
                orient ? orient[1] : 2, orient ? orient[2] : 3);
      return false;
    }
    const int a = (code - 1) % 3;
    const bool back = code > 3;
    seen[a] = true;
    n[a] = ext[m];
    lo[a] = back ? s[a] - (ext[m] - 1) : s[a];
    p.step[m] = back ? -stride[a] : stride[a];
  }
  const pga_box bs{{lo[0], lo[1], lo[2]}, {n[0], n[1], n[2]}};
  if (!covers(src, bs, ncomp)) {
    set_error("%s: the source cells reach outside their array", who);
    return false;
  }
  p.base = s[0] + s[1] * src.sj + s[2] * src.sk;
  return true;
}

// A rotational periodic's turn of the vectors that cross it (place.mako:3-37): R row-major; a
// task turns only when R is given and is not exactly the identity, so an unturned copy stays a copy
// to the bit
struct Turn {
  double R[9];
  int on;
};

inline Turn turn_of(const double *R) {
  Turn t{};
  if (!R) return t;
  for (int n = 0; n < 9; n++) {
    t.R[n] = R[n];
    if (R[n] != (n % 4 == 0 ? 1.0 : 0.0)) t.on = 1;
  }
  return t;
}

#define PGA_REQUIRE(cond, ...)      \
  do {                              \
    if (!(cond)) {                  \
      set_error(__VA_ARGS__);       \
      return PGA_ERANGE;            \
    }                               \
  } while (0)

// a library built for one case (peregrine_amd/build.py, build_variant) names its own species counts
// and real-gas counts in generated/variant.hpp; the main library's are these
#if __has_include("generated/variant.hpp")
#include "generated/variant.hpp"
#endif
#ifndef PGA_NS_CASES
#define PGA_NS_CASES(X) X(1) X(2) X(3) X(9)
#endif
#ifndef PGA_REALGAS_CASES
#define PGA_REALGAS_CASES(ns) ((ns) == 1 || (ns) == 2)
#endif

// doubles of a cell's cache of the implicit chemistry's f at one T for the case's mechanism, -1 for
// none (pga_implicit.hip)
int64_t im_cache_doubles(const pga_case_desc &d);

// the species counts the cubic real gas is compiled for (its kernels are the heaviest to build)
constexpr bool realgas_compiled(int ns) { return PGA_REALGAS_CASES(ns); }

// the state fields a case stores: the real gas adds its compressibility Z
inline int nstate(const pga_case_desc &d) { return d.eos == 2 ? 9 : 8; }

// f.template operator()<NS, EOS>() for the case's species count and eos (0 cpg, 1 tpg, 2 the
// cubic real gas where compiled); false when no variant is compiled for them
template <class F>
bool dispatch_eos(const pga_case_desc &d, F &&f) {
  bool done = false;
  const auto one = [&]<int N>() {
    if (d.ns != N) return;
    if (d.eos == 0) {
      f.template operator()<N, 0>();
      done = true;
    } else if (d.eos == 1) {
      f.template operator()<N, 1>();
      done = true;
    } else if constexpr (realgas_compiled(N)) {
      if (d.eos == 2) {
        f.template operator()<N, 2>();
        done = true;
      }
    }
  };
#define X(N) one.template operator()<N>();
  PGA_NS_CASES(X)
#undef X
  return done;
}

}  // namespace pga
