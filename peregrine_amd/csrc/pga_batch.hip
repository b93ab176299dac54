// Launches over a table of blocks: a solver with many blocks (an uneven batch of independent boxes,
// or a block lattice) runs each phase of its fused stage as ONE launch over every block, where the
// block-at-a-time entry points make one launch per block and phase. The reference tiles every block
// of a rank into one launch the same way (launch/table.py:83-106, a range table per kernel); here the
// table lives in a device buffer the caller owns, built once (pga_batch_create), and its pointers
// (the blocks' registers, state, faces and geometry) stay valid for the solver's life, so the
// launches are graph-capturable as they are.
//
// Each launch's workgroups are numbered across the blocks by a prefix table of tiles per block; a
// workgroup finds its block by a binary search (wave-uniform, so scalar loads) and its tile within
// the block, then runs the per-cell work the block-at-a-time kernels run (pga_stagecore.hpp,
// flux_face), so a block's result is the same to the bit. The XCD-ordered workgroup numbering walks
// each XCD through one contiguous run of the blocks' tiles.
//
// Tile shapes are per block, so one launch can give a 16-cell-wide block 16-wide tiles and a
// 96-wide one 64-wide tiles: the stage's column tiles (wi x 256/wi columns, kc planes per thread)
// and the k-sweep's (wi faces along i x 256/wi along k) are chosen per block-size class by the
// project's autotuner (peregrine_amd/autotune.py), which times the candidates on the case itself.
#include <cstring>
#include <vector>

#include "pga_implicit.hpp"
#include "pga_internal.hpp"
#include "pga_stagecore.hpp"

namespace pga {

enum { BK_CELLS = 0, BK_FLUX0, BK_FLUX1, BK_FLUX2, BK_STAGE, BK_N };

struct BBlock {
  View Q[4], st, S[3], F[3], det, rates;
  Box3 b;      // the interior
  int wi, kc;  // the stage's tile width along i and the planes a thread walks
  int ki;      // the k-sweep's tile width along i
  int pad;
};

struct BSlab {
  int blk, src;  // the block whose halo this is, the block it reads
  int d0, d1, d2, e0, e1;
  int64_t sbase, st0, st1, st2;  // the source cell of the slab's first cell, the steps of its axes
};

}  // namespace pga

struct pga_batch {
  const pga_case *c;
  int nblocks, nslabs;
  pga::BBlock *blocks;           // device
  pga::BSlab *slabs;             // device
  int64_t *prefix[pga::BK_N];    // device: nblocks + 1 tile starts per kind
  int64_t *slabprefix;           // device: nslabs + 1 cell starts
  int64_t total[pga::BK_N];      // host copies of the totals
  int64_t slabcells;
  int32_t regs;                  // registers each block holds
};

namespace pga {

__device__ __forceinline__ int bfind(const int64_t *__restrict__ pre, int n, int64_t g) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct BArgs {
  const double *K, *C;
  const BBlock *blocks;
  const BSlab *slabs;
  const int64_t *prefix, *slabprefix;
  int nblocks, nslabs;
  int q, qout, axis;
};

template <class M>
__global__ __launch_bounds__(FLAT) void kb_rates(BArgs a) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (!flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  rates_cell<M>(a.K, a.C, B.Q[a.q], B.st, B.rates, i, j, k);
}

template <int NS, int EOS>
__global__ __launch_bounds__(FLAT) void kb_eos(BArgs a) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (!flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  eos_cell<NS, EOS>(a.K, B.Q[a.qout], B.st, B.st, i, j, k);
}

// the halo pass over every block's slabs: one thread per halo cell, the slabs back to back
template <int NS, int EOS>
__global__ __launch_bounds__(FLAT) void kb_halo(BArgs a) {
  const int64_t c = (int64_t)blockIdx.x * FLAT + threadIdx.x;
  if (c >= a.slabprefix[a.nslabs]) return;
  const int n = bfind(a.slabprefix, a.nslabs, c);
  const BSlab &t = a.slabs[n];
  const int64_t r = c - a.slabprefix[n];
  const int ii = (int)(r % t.e0), jj = (int)((r / t.e0) % t.e1), kk = (int)(r / ((int64_t)t.e0 * t.e1));
  const int64_t os = t.sbase + ii * t.st0 + jj * t.st1 + kk * t.st2;
  const BBlock &D = a.blocks[t.blk];
  const Turn none{};
  halo_cell<NS, EOS>(a.K, D.Q[a.q], a.blocks[t.src].Q[a.q], os, none, D.st, D.st, t.d0 + ii, t.d1 + jj, t.d2 + kk);
}

// one axis's faces of every block: i and j faces in flat tiles, k faces in wi x (256 / wi) tiles
template <int NS, int EOS, int RECON, int PRIMARY, int LIM>
__global__ __launch_bounds__(FLAT) void kb_flux(BArgs a) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  const int64_t t = g - a.prefix[n];
  Box3 fb = B.b;
  int i, j, k;
  if (a.axis == 2) {
    fb.e2 += 1;
    // a column's k tiles one after another, as k_faces takes them
    const int wi = B.ki, kt = FLAT / wi;
    const int64_t nbi = (fb.e0 + wi - 1) / wi, rest = t / nbi, nkt = (fb.e2 + kt - 1) / kt;
    i = fb.s0 + (int)(t % nbi) * wi + (threadIdx.x & (wi - 1));
    k = fb.s2 + (int)(rest % nkt) * kt + threadIdx.x / wi;
    j = fb.s1 + (int)(rest / nkt);
    if (i >= fb.s0 + fb.e0 || k >= fb.s2 + fb.e2) return;
  } else {
    if (a.axis == 0)
      fb.e0 += 1;
    else
      fb.e1 += 1;
    if (!flat_at(fb, t, i, j, k)) return;
  }
  flux_face<NS, EOS, RECON, PRIMARY, LIM>(a.K, a.axis, B.Q[a.q], B.st, B.S[a.axis], B.F[a.axis], i, j, k);
}

struct BStage {
  CombineSpec comb;
  const double *dt;
  int q0;
};

// the divergence, the source and the combine of every block's cells, in wi x (256 / wi) column tiles
template <int NE, bool CHEM>
__global__ __launch_bounds__(FLAT) void kb_stage(BArgs a, BStage s) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  const int64_t t = g - a.prefix[n];
  const int wi = B.wi, wj = FLAT / wi;
  const int64_t nbi = (B.b.e0 + wi - 1) / wi, nbj = (B.b.e1 + wj - 1) / wj;
  const int i = B.b.s0 + (int)(t % nbi) * wi + (threadIdx.x & (wi - 1));
  const int j = B.b.s1 + (int)((t / nbi) % nbj) * wj + threadIdx.x / wi;
  if (i >= B.b.s0 + B.b.e0 || j >= B.b.s1 + B.b.e1) return;
  const int kbeg = B.b.s2 + (int)(t / (nbi * nbj)) * B.kc;
  StageArgs sa{};
  sa.Qin = B.Q[a.q], sa.Q0 = B.Q[0], sa.Qout = B.Q[a.qout], sa.det = B.det, sa.rates = B.rates;
  sa.iF = B.F[0], sa.jF = B.F[1], sa.kF = B.F[2];
  sa.comb = s.comb, sa.dt = s.dt, sa.q0 = s.q0;
  stage_cells<NE, CHEM>(sa, i, j, kbeg, min(kbeg + B.kc, B.b.s2 + B.b.e2));
}

// implicit chemistry over every block's cells into their rates: a cell's place in the work arrays is
// its workgroup's flat number times 256 plus its lane, the same in both launches
template <class M>
__global__ __launch_bounds__(FLAT) void kb_itc(BArgs a, ImArgs im) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (!flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  ijac_tcache<M>(im, B.Q[a.q], B.st, i, j, k, g * FLAT + threadIdx.x);
}

template <class M>
__global__ __launch_bounds__(FLAT) void kb_ibase(BArgs a, ImArgs im) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (!flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  ijac_base<M>(im, B.Q[a.q], i, j, k, g * FLAT + threadIdx.x);
}

// a tile's slices (the ns + 1 columns and f(y)) in workgroups numbered one after another, as k_icol
template <class M>
__global__ __launch_bounds__(FLAT) void kb_icol(BArgs a, ImArgs im) {
  constexpr int64_t S = M::NS + 2;
  const int64_t gs = xcd_block(), g = gs / S;
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (!flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  ijac_column<M>(im, B.Q[a.q], i, j, k, g * FLAT + threadIdx.x, (int)(gs - g * S));
}

// the places of a tile's overhang (no cell): the identity and a zero right-hand side, once a call,
// so the quad-per-cell solve finds every place of its groups defined and settles them in one sweep
template <int NV>
__global__ __launch_bounds__(FLAT) void kb_ivoid(BArgs a, ImArgs im) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  for (int m = 0; m <= NV; m++) ijac_void<NV>(im, g * FLAT + threadIdx.x, m);
}

template <class M>
__global__ __launch_bounds__(FLAT) void kb_iend(BArgs a, ImArgs im) {
  const int64_t g = xcd_block();
  const int n = bfind(a.prefix, a.nblocks, g);
  const BBlock &B = a.blocks[n];
  int i, j, k;
  if (!flat_at(B.b, g - a.prefix[n], i, j, k)) return;
  iend_cell<M>(im, B.Q[a.q], B.rates, i, j, k, g * FLAT + threadIdx.x);
}

// the tiles of each kind a block takes
static void block_tiles(const BBlock &B, int64_t out[BK_N]) {
  const Box3 &b = B.b;
  const int64_t cells = (int64_t)b.e0 * b.e1 * b.e2;
  out[BK_CELLS] = (cells + FLAT - 1) / FLAT;
  out[BK_FLUX0] = ((int64_t)(b.e0 + 1) * b.e1 * b.e2 + FLAT - 1) / FLAT;
  out[BK_FLUX1] = ((int64_t)b.e0 * (b.e1 + 1) * b.e2 + FLAT - 1) / FLAT;
  const int kt = FLAT / B.ki;
  out[BK_FLUX2] = (int64_t)((b.e0 + B.ki - 1) / B.ki) * b.e1 * ((b.e2 + 1 + kt - 1) / kt);
  const int wj = FLAT / B.wi;
  out[BK_STAGE] = (int64_t)((b.e0 + B.wi - 1) / B.wi) * ((b.e1 + wj - 1) / wj) * ((b.e2 + B.kc - 1) / B.kc);
  if (!cells)
    for (int k = 0; k < BK_N; k++) out[k] = 0;
}

static bool tile_width(int w) { return w == 16 || w == 32 || w == 64; }

// scheme(f): f.template operator()<RECON, PRIMARY, LIM>() for the case's face scheme
template <class F>
static bool dispatch_scheme(const pga_case_desc &d, F &&f) {
  bool done = false;
  const auto one = [&]<int RECON, int PRIM, int LIM>() {
    if (d.recon == RECON && d.primary == PRIM && (RECON == RECON_CONSTANT || d.limiter == LIM)) {
      f.template operator()<RECON, PRIM, LIM>();
      done = true;
    }
  };
  one.template operator()<RECON_CONSTANT, FLUX_KEPAEC, LIM_NONE>();
  one.template operator()<RECON_CONSTANT, FLUX_HLLC, LIM_NONE>();
  one.template operator()<RECON_CONSTANT, FLUX_RUSANOV, LIM_NONE>();
#define M(PRIM)                                                   \
  one.template operator()<RECON_MUSCL, PRIM, LIM_KOREN>();        \
  one.template operator()<RECON_MUSCL, PRIM, LIM_MINMOD>();       \
  one.template operator()<RECON_MUSCL, PRIM, LIM_MC>();           \
  one.template operator()<RECON_MUSCL, PRIM, LIM_SUPERBEE>();     \
  one.template operator()<RECON_MUSCL, PRIM, LIM_VANLEER>();
  M(FLUX_HLLC)
  M(FLUX_RUSANOV)
  M(FLUX_KEPAEC)
#undef M
  return done;
}

static BArgs bargs(const pga_batch *h, int q, int qout, int axis, const int64_t *prefix) {
  BArgs a{};
  a.K = h->c->K, a.C = h->c->chem;
  a.blocks = h->blocks, a.slabs = h->slabs, a.prefix = prefix, a.slabprefix = h->slabprefix;
  a.nblocks = h->nblocks, a.nslabs = h->nslabs;
  a.q = q, a.qout = qout, a.axis = axis;
  return a;
}

static dim3 tiles_grid(int64_t n) { return dim3((unsigned)n); }

// bytes of the table for n blocks and m slabs, each part 256-byte aligned
static int64_t table_layout(int n, int m, int64_t off[4]) {
  const auto up = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  off[0] = 0;
  off[1] = up(off[0] + (int64_t)n * sizeof(BBlock));
  off[2] = up(off[1] + (int64_t)m * sizeof(BSlab));
  off[3] = up(off[2] + (int64_t)BK_N * (n + 1) * sizeof(int64_t));
  return up(off[3] + (int64_t)(m + 1) * sizeof(int64_t));
}

}  // namespace pga

using namespace pga;

extern "C" int64_t pga_batch_bytes(int32_t nblocks, int32_t nslabs) {
  if (nblocks < 0 || nslabs < 0) return -1;
  int64_t off[4];
  return table_layout(nblocks, nslabs, off);
}

extern "C" int32_t pga_batch_create(const pga_case *c, int32_t nblocks, const pga_batch_block *blocks, int32_t nregs,
                                    void *table, int64_t nbytes, pga_batch **out) {
  if (!c || nblocks < 1 || !blocks || nregs < 1 || nregs > 4 || !table || !out) {
    set_error("pga_batch_create: a case, one or more blocks of 1 to 4 registers, a table and an out pointer");
    return PGA_EINVAL;
  }
  const pga_case_desc &d = c->desc;
  const int ne = d.ns + 4, ns = nstate(d);
  const bool chem = d.mechanism >= 0;
  int nslabs = 0;
  std::vector<BBlock> bb(nblocks);
  for (int n = 0; n < nblocks; n++) {
    const pga_batch_block &s = blocks[n];
    const pga_box &box = s.interior;
    PGA_REQUIRE(box_ok(box) && !empty(box) && flat_fits(box), "pga_batch_create: block %d has a bad or empty interior",
                n);
    PGA_REQUIRE(s.nslab >= 0 && s.nslab <= 6, "pga_batch_create: block %d has %d halo slabs (0 to 6)", n, s.nslab);
    nslabs += s.nslab;
    // what every phase reads and writes: the stage's, the sweeps' (MUSCL's stencil of two cells each
    // side) and the halo pass's own checks below
    const int z[3] = {0, 0, 0};
    int lo[3], hi[3];
    for (int r = 0; r < nregs; r++) {
      PGA_REQUIRE(covers(s.Q[r], box, ne), "pga_batch_create: block %d's register %d does not cover its interior", n, r);
      PGA_REQUIRE(s.Q[r].sj == s.Q[0].sj && s.Q[r].sk == s.Q[0].sk,
                  "pga_batch_create: block %d's registers differ in layout", n);
    }
    PGA_REQUIRE(covers(s.state, box, ns) && covers(s.detJinv, box, 1),
                "pga_batch_create: block %d's state or detJinv does not cover its interior", n);
    const pga_view *S[3] = {&s.iS, &s.jS, &s.kS}, *F[3] = {&s.iF, &s.jF, &s.kF};
    for (int a = 0; a < 3; a++) {
      for (int m = 0; m < 3; m++) lo[m] = hi[m] = 0;
      lo[a] = d.recon == 1 ? 2 : 1;
      hi[a] = d.recon == 1 ? 2 : 1;
      for (int r = 0; r < nregs; r++)
        PGA_REQUIRE(covers(s.Q[r], box, lo, hi, ne), "pga_batch_create: block %d's faces' stencil leaves register %d",
                    n, r);
      PGA_REQUIRE(covers(s.state, box, lo, hi, ns), "pga_batch_create: block %d's faces' stencil leaves its state", n);
      int fh[3] = {0, 0, 0};
      fh[a] = 1;
      PGA_REQUIRE(covers(*S[a], box, z, fh, 3) && covers(*F[a], box, z, fh, ne),
                  "pga_batch_create: block %d's faces reach outside S or F along axis %d", n, a);
    }
    if (chem)
      PGA_REQUIRE(covers(s.rates, box, d.ns - 1), "pga_batch_create: block %d has no rates array", n);
    const int wi = s.stage_wi ? s.stage_wi : 64, kc = s.stage_kc ? s.stage_kc : 1, ki = s.flux_ki ? s.flux_ki : 64;
    PGA_REQUIRE(tile_width(wi) && tile_width(ki) && kc >= 1 && kc <= 64,
                "pga_batch_create: block %d's tiles: widths 16, 32 or 64 and 1 to 64 planes", n);
    BBlock &B = bb[n];
    for (int r = 0; r < nregs; r++) B.Q[r] = view(s.Q[r]);
    B.st = view(s.state), B.det = view(s.detJinv), B.rates = view(s.rates);
    for (int a = 0; a < 3; a++) B.S[a] = view(*S[a]), B.F[a] = view(*F[a]);
    B.b = box3(box), B.wi = wi, B.kc = kc, B.ki = ki;
  }
  std::vector<BSlab> sl;
  sl.reserve(nslabs);
  for (int n = 0; n < nblocks; n++) {
    const pga_batch_block &s = blocks[n];
    for (int t = 0; t < s.nslab; t++) {
      const int src = s.slab_src[t];
      PGA_REQUIRE(src >= 0 && src < nblocks, "pga_batch_create: block %d's slab %d reads block %d of %d", n, t, src,
                  nblocks);
      const int *ds = s.slab_dst[t], *e = s.slab_extent[t];
      const pga_box bd{{ds[0], ds[1], ds[2]}, {e[0], e[1], e[2]}};
      PGA_REQUIRE(box_ok(bd), "pga_batch_create: block %d's slab %d is a bad box", n, t);
      for (int r = 0; r < nregs; r++)
        PGA_REQUIRE(covers(s.Q[r], bd, ne), "pga_batch_create: block %d's slab %d leaves register %d", n, t, r);
      PGA_REQUIRE(covers(s.state, bd, ns), "pga_batch_create: block %d's slab %d leaves its state", n, t);
      PGA_REQUIRE(s.Q[0].sj == s.state.sj && s.Q[0].sk == s.state.sk,
                  "pga_batch_create: block %d's Q and state must share their layout", n);
      for (int m = 0; m < 3; m++)
        PGA_REQUIRE(s.slab_from[t][m] >= 0 && e[m] >= 0, "pga_batch_create: a negative source start or extent");
      Placement pl;
      for (int r = 0; r < nregs; r++)
        if (!placement("pga_batch_create", blocks[src].Q[r], s.slab_from[t], s.slab_orient[t], e, ne, pl))
          return PGA_ERANGE;
      BSlab b{};
      b.blk = n, b.src = src, b.d0 = ds[0], b.d1 = ds[1], b.d2 = ds[2], b.e0 = e[0], b.e1 = e[1];
      b.sbase = pl.base, b.st0 = pl.step[0], b.st1 = pl.step[1], b.st2 = pl.step[2];
      sl.push_back(b);
    }
  }
  int64_t off[4];
  const int64_t need = table_layout(nblocks, nslabs, off);
  PGA_REQUIRE(nbytes >= need, "pga_batch_create: the table holds %lld bytes, %d blocks and %d slabs need %lld",
              (long long)nbytes, nblocks, nslabs, (long long)need);
  std::vector<char> img(need, 0);
  std::memcpy(img.data() + off[0], bb.data(), bb.size() * sizeof(BBlock));
  if (nslabs) std::memcpy(img.data() + off[1], sl.data(), sl.size() * sizeof(BSlab));
  int64_t *pre = reinterpret_cast<int64_t *>(img.data() + off[2]);
  pga_batch h{};
  for (int k = 0; k < BK_N; k++) pre[k * (nblocks + 1)] = 0;
  for (int n = 0; n < nblocks; n++) {
    int64_t t[BK_N];
    block_tiles(bb[n], t);
    for (int k = 0; k < BK_N; k++) pre[k * (nblocks + 1) + n + 1] = pre[k * (nblocks + 1) + n] + t[k];
  }
  for (int k = 0; k < BK_N; k++) {
    h.total[k] = pre[k * (nblocks + 1) + nblocks];
    PGA_REQUIRE(h.total[k] < (int64_t(1) << 31), "pga_batch_create: %lld workgroups in one launch",
                (long long)h.total[k]);
  }
  int64_t *sp = reinterpret_cast<int64_t *>(img.data() + off[3]);
  sp[0] = 0;
  // each slab's cells
  {
    int t = 0;
    for (int n = 0; n < nblocks; n++)
      for (int m = 0; m < blocks[n].nslab; m++, t++)
        sp[t + 1] = sp[t] + (int64_t)blocks[n].slab_extent[m][0] * blocks[n].slab_extent[m][1] *
                                blocks[n].slab_extent[m][2];
  }
  h.slabcells = sp[nslabs];
  if (hipMemcpy(table, img.data(), need, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("pga_batch_create: the table's upload failed");
    return PGA_EHIP;
  }
  char *base = static_cast<char *>(table);
  h.c = c, h.nblocks = nblocks, h.nslabs = nslabs, h.regs = nregs;
  h.blocks = reinterpret_cast<BBlock *>(base + off[0]);
  h.slabs = reinterpret_cast<BSlab *>(base + off[1]);
  for (int k = 0; k < BK_N; k++) h.prefix[k] = reinterpret_cast<int64_t *>(base + off[2]) + k * (nblocks + 1);
  h.slabprefix = reinterpret_cast<int64_t *>(base + off[3]);
  *out = new pga_batch(h);
  return PGA_OK;
}

extern "C" int32_t pga_batch_destroy(pga_batch *h) {
  delete h;
  return PGA_OK;
}

extern "C" int32_t pga_batch_halo(const pga_batch *h, int32_t q, void *stream) {
  if (!h || q < 0 || q >= h->regs) {
    set_error("pga_batch_halo: a batch and one of its registers");
    return PGA_EINVAL;
  }
  if (!h->slabcells) return PGA_OK;
  const BArgs a = bargs(h, q, q, 0, h->prefix[BK_CELLS]);
  const dim3 grid((unsigned)((h->slabcells + FLAT - 1) / FLAT));
  const bool done = dispatch_eos(h->c->desc, [&]<int N, int E>() {
    hipLaunchKernelGGL((kb_halo<N, E>), grid, dim3(FLAT), 0, stream_of(stream), a);
  });
  if (!done) {
    set_error("pga_batch_halo: no kernel for %d species with eos %d", h->c->desc.ns, h->c->desc.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_batch_halo");
}

extern "C" int32_t pga_batch_rates(const pga_batch *h, int32_t q, void *stream) {
  if (!h || q < 0 || q >= h->regs) {
    set_error("pga_batch_rates: a batch and one of its registers");
    return PGA_EINVAL;
  }
  if (h->c->desc.mechanism < 0 || !h->c->chem) {
    set_error("pga_batch_rates: the case has no compiled mechanism");
    return PGA_EUNSUP;
  }
  const BArgs a = bargs(h, q, q, 0, h->prefix[BK_CELLS]);
  bool done = false;
#define X(NAME, IDX)                                                                                       \
  if (h->c->desc.mechanism == IDX) {                                                                       \
    hipLaunchKernelGGL((kb_rates<Mech_##NAME>), tiles_grid(h->total[BK_CELLS]), dim3(FLAT), 0,           \
                       stream_of(stream), a);                                                              \
    done = true;                                                                                           \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_batch_rates: no mechanism %d", h->c->desc.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_batch_rates");
}

extern "C" int32_t pga_batch_flux(const pga_batch *h, int32_t q, int32_t axis, void *stream) {
  if (!h || q < 0 || q >= h->regs || axis < 0 || axis > 2) {
    set_error("pga_batch_flux: a batch, one of its registers and an axis");
    return PGA_EINVAL;
  }
  const pga_case_desc &d = h->c->desc;
  const BArgs a = bargs(h, q, q, axis, h->prefix[BK_FLUX0 + axis]);
  const dim3 grid = tiles_grid(h->total[BK_FLUX0 + axis]);
  bool done = false;
  dispatch_eos(d, [&]<int N, int E>() {
    done = dispatch_scheme(d, [&]<int RECON, int PRIM, int LIM>() {
      hipLaunchKernelGGL((kb_flux<N, E, RECON, PRIM, LIM>), grid, dim3(FLAT), 0, stream_of(stream), a);
    });
  });
  if (!done) {
    set_error("pga_batch_flux: no kernel for %d species, eos %d, reconstruction %d, flux %d, limiter %d", d.ns, d.eos,
              d.recon, d.primary, d.limiter);
    return PGA_EUNSUP;
  }
  return launched("pga_batch_flux");
}

extern "C" int32_t pga_batch_stage(const pga_batch *h, int32_t q, int32_t qout, int32_t nterms, const int32_t *src,
                                   const double *w, const int32_t *withdt, const double *dt, void *stream) {
  if (!h || q < 0 || q >= h->regs || qout < 0 || qout >= h->regs || !src || !w || !withdt || !dt || nterms < 1 ||
      nterms > 4) {
    set_error("pga_batch_stage: a batch, two of its registers, 1 to 4 combine terms and dt");
    return PGA_EINVAL;
  }
  BStage s{};
  s.comb.n = nterms;
  for (int t = 0; t < nterms; t++) {
    PGA_REQUIRE(src[t] >= 0 && src[t] <= 2, "pga_batch_stage: a combine term reads Qin (0), Q0 (1) or dQ (2)");
    s.comb.src[t] = src[t], s.comb.w[t] = w[t], s.comb.withdt[t] = withdt[t];
    s.q0 |= src[t] == SRC_Q0;
  }
  s.dt = dt;
  const pga_case_desc &d = h->c->desc;
  const bool chem = d.mechanism >= 0;
  const BArgs as = bargs(h, q, qout, 0, h->prefix[BK_STAGE]), ae = bargs(h, q, qout, 0, h->prefix[BK_CELLS]);
  const bool done = dispatch_eos(d, [&]<int N, int E>() {
    if (chem)
      hipLaunchKernelGGL((kb_stage<N + 4, true>), tiles_grid(h->total[BK_STAGE]), dim3(FLAT), 0, stream_of(stream),
                         as, s);
    else
      hipLaunchKernelGGL((kb_stage<N + 4, false>), tiles_grid(h->total[BK_STAGE]), dim3(FLAT), 0, stream_of(stream),
                         as, s);
    hipLaunchKernelGGL((kb_eos<N, E>), tiles_grid(h->total[BK_CELLS]), dim3(FLAT), 0, stream_of(stream), ae);
  });
  if (!done) {
    set_error("pga_batch_stage: no kernel for %d species with eos %d", d.ns, d.eos);
    return PGA_EUNSUP;
  }
  return launched("pga_batch_stage");
}

// the cells the implicit form's work arrays hold for a batch: every cell tile whole
static int64_t batch_cells(const pga_batch *h) { return h->total[BK_CELLS] * FLAT; }

extern "C" int64_t pga_batch_implicit_worksize(const pga_batch *h) {
  if (!h || h->c->desc.ns + 1 > IM_N) return -1;
  const int64_t cells = batch_cells(h), groups = (cells + IM_G - 1) / IM_G, nv = h->c->desc.ns + 1;
  const int64_t nt = im_cache_doubles(h->c->desc);
  if (nt < 0) return -1;
  return groups * nv * nv * IM_G + (3 * nv + nt) * cells;
}

extern "C" int32_t pga_batch_implicit(const pga_batch *h, const pga_implicit_opts *o, int32_t q, int32_t *sweeps,
                                      int64_t nsweeps, double *work, int64_t nwork, const double *dt, void *stream) {
  if (!h || !o || q < 0 || q >= h->regs || !sweeps || !work || !dt) {
    set_error("pga_batch_implicit: a batch, options, one of its registers, sweeps, work and dt");
    return PGA_EINVAL;
  }
  if (o->parts < 1 || o->max_sweeps < 1 || !(o->tolerance > 0.0)) {
    set_error("pga_batch_implicit: parts >= 1, max_sweeps >= 1 and a tolerance above zero are required");
    return PGA_EINVAL;
  }
  const pga_case_desc &d = h->c->desc;
  if (d.mechanism < 0 || !h->c->chem || d.ns + 1 > IM_N) {
    set_error("pga_batch_implicit: the case has no compiled mechanism, or more species than the LU takes");
    return PGA_EUNSUP;
  }
  const int64_t cells = batch_cells(h), groups = (cells + IM_G - 1) / IM_G, nv = d.ns + 1;
  PGA_REQUIRE(nsweeps >= cells, "pga_batch_implicit: sweeps holds %lld counts, the batch needs %lld",
              (long long)nsweeps, (long long)cells);
  PGA_REQUIRE(nwork >= pga_batch_implicit_worksize(h), "pga_batch_implicit: work holds %lld doubles, needs %lld",
              (long long)nwork, (long long)pga_batch_implicit_worksize(h));
  ImArgs im{};
  im.K = h->c->K, im.C = h->c->chem;
  im.begin_dQ = 0;
  im.A = work;
  im.y = im.A + groups * nv * nv * IM_G;
  im.tc = im.y + 3 * nv * cells;
  im.ncell = cells, im.sweeps = sweeps, im.dt = dt;
  im.parts = o->parts, im.maxsweeps = o->max_sweeps, im.tol = o->tolerance;
  const BArgs a = bargs(h, q, q, 0, h->prefix[BK_CELLS]);
  const dim3 grid = tiles_grid(h->total[BK_CELLS]);
  hipStream_t s = stream_of(stream);
  bool done = false;
#define X(NAME, IDX)                                                                             \
  if (d.mechanism == IDX) {                                                                      \
    constexpr int NV = Mech_##NAME::NS + 1, NB = (NV + 3) / 4;                                   \
    hipLaunchKernelGGL((kb_ivoid<NV>), grid, dim3(FLAT), 0, s, a, im);                           \
    for (int part = 0; part < o->parts; part++) {                                                \
      im.first = part == 0, im.last = part == o->parts - 1;                                      \
      hipLaunchKernelGGL((kb_itc<Mech_##NAME>), grid, dim3(FLAT), 0, s, a, im);                 \
      hipLaunchKernelGGL((kb_ibase<Mech_##NAME>), grid, dim3(FLAT), 0, s, a, im);               \
      hipLaunchKernelGGL((kb_icol<Mech_##NAME>), dim3(grid.x * (NV + 1)), dim3(FLAT), 0, s, a, im); \
      hipLaunchKernelGGL((k_lusolve<NB, NV>), dim3((unsigned)((groups + 3) / 4)), dim3(256), 0, s, im); \
    }                                                                                            \
    hipLaunchKernelGGL((kb_iend<Mech_##NAME>), grid, dim3(FLAT), 0, s, a, im);                  \
    done = true;                                                                                 \
  }
  PGA_MECHANISMS(X)
#undef X
  if (!done) {
    set_error("pga_batch_implicit: no mechanism %d", d.mechanism);
    return PGA_EUNSUP;
  }
  return launched("pga_batch_implicit");
}
