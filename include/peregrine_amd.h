/*
 * peregrine_amd: the C ABI of an MI355X (gfx950) implementation of Peregrine's
 * per-Runge-Kutta-stage right-hand side and stage update.
 *
 * Plain pointers and sizes only. Every array is device memory owned by the
 * caller, in the reference's device layout: Fortran order [i, j, k, comp],
 * i contiguous, each component a separate plane, ng halo layers on every side
 * (peregrine/backend/device/backend.py:56-57, peregrine/launch/ranges.py:19-20).
 * A face array along an axis has one plane more than the cells along it; face
 * index f lies between cells f-1 and f (peregrine/launch/policy.py:76-93).
 *
 * Every entry point returns 0 on success or a negative PGA_E* code; the text
 * of the last error on the calling thread is pga_last_error(). Launches are
 * asynchronous on the HIP stream passed (NULL: the default stream).
 *
 * The reference binds one entry point per rendered kernel,
 *   extern "C" __global__ void pg_<symbol>(const Functor k, const pg::Tiling t)
 * (peregrine/backend/device/kernels/launch.mako:1-12); each function below
 * names the template whose pg_<symbol> it replaces.
 */
#ifndef PEREGRINE_AMD_H
#define PEREGRINE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGA_ABI_VERSION 4

enum {
  PGA_OK = 0,
  PGA_EINVAL = -1,   /* an argument the kernels cannot take */
  PGA_EUNSUP = -2,   /* a case this build has no kernel for (species count, mechanism) */
  PGA_EHIP = -3,     /* the HIP runtime refused a launch */
  PGA_ERANGE = -4,   /* a size beyond a stated maximum */
};

/* one array: base pointer at element (0,0,0,0), its j, k and component strides in elements, and
 * its extents (i, j, k, components); every entry point checks its boxes and stencils against the
 * extents before it launches */
typedef struct {
  double *data;
  int64_t sj, sk, sc;
  int32_t n[4];
} pga_view;

/* a box of array indices: [start, start + extent) per axis, halo cells included in the index */
typedef struct {
  int32_t start[3];
  int32_t extent[3];
} pga_box;

/* the kernel variants of a case, from its config (peregrine/config.py:12-98) */
typedef struct {
  int32_t ns;        /* species; ne = 4 + ns conserved variables */
  int32_t eos;       /* 0 cpg, 1 tpg */
  int32_t recon;     /* 0 constant, 1 muscl */
  int32_t limiter;   /* 0 none, 1 koren, 2 minmod, 3 mc, 4 superbee, 5 vanleer */
  int32_t primary;   /* 0 kepaec, 1 hllc, 2 rusanov */
  int32_t mechanism; /* -1 none, else the index of a compiled mechanism (pga_mechanism_find) */
  int32_t trans;     /* -1 none (Euler), 0 constant, 1 kinetic theory, 2 Chung; Wilke mixing */
  int32_t diffusion; /* 0 Lewis numbers, 1 binary (mixture-averaged pair coefficients) */
  int32_t reserved[8];
} pga_case_desc;

typedef struct pga_case pga_case;

/* library and error reporting */
int32_t pga_abi_version(void);
const char *pga_last_error(void);

/* The case constant table: per species MW, MWinv, Rsp, cppolyR[7], h298R, cp0, gpoly[8], h298, ...
 * in one flat double array on the device. The offset and length of each named section for ns
 * species; -1 when the name is unknown. Mirrors the constants the reference bakes into every unit
 * (peregrine/simulator/mixture/mixture.py:54-70, peregrine/codegen/constants.py:7-65). */
int32_t pga_const_offset(int32_t ns, const char *name, int32_t *count);
int32_t pga_const_size(int32_t ns);

/* compiled finite-rate mechanisms: count, the structure signature of one, and the length of its
 * constant table; pga_mechanism_find returns the index whose signature equals the string, or -1 */
int32_t pga_mechanism_count(void);
const char *pga_mechanism_signature(int32_t index);
const char *pga_mechanism_name(int32_t index);
int32_t pga_mechanism_table_size(int32_t index);
int32_t pga_mechanism_find(const char *signature);

/* a case: the variant choice plus the device tables it reads (borrowed, not copied or freed) */
int32_t pga_case_create(const pga_case_desc *desc, const double *dev_consts, const double *dev_chem,
                        pga_case **out);
int32_t pga_case_destroy(pga_case *c);

/* EOS inversion Q -> state (p, T, rhoinv, rho, e, cp, gamma, c) over up to 8 boxes in one launch;
 * state's T is read as the Newton start.
 * Replaces pg_fromQ: peregrine/simulator/euler/kernels/state/fromQ.mako:5-19. */
int32_t pga_fromQ(const pga_case *c, pga_view Q, pga_view state, const pga_box *boxes, int32_t nboxes,
                  void *stream);

/* primitives [p, u, v, w, T, Y_0..Y_{ns-2}] -> Q and state.
 * Replaces pg_fromprims: peregrine/simulator/euler/kernels/state/fromprims.mako:4-13. */
int32_t pga_fromprims(const pga_case *c, pga_view prims, pga_view Q, pga_view state, pga_box box, void *stream);

/* the advective flux F (ne per face) through the faces of one axis in box (face indices).
 * Replaces pg_flux: peregrine/simulator/euler/kernels/flux/single.mako:4-9 with recon/{constant,muscl}
 * and function/{kepaec,hllc,rusanov}. */
int32_t pga_flux(const pga_case *c, int32_t axis, pga_view Q, pga_view state, pga_view S, pga_view F,
                 pga_box box, void *stream);

/* Every face of a block in one launch: for each axis a (0 i, 1 j, 2 k), the faces boxes[a] of S[a]
 * and F[a], each face's advective flux as pga_flux gives it and, when viscous, the alpha-damped viscous
 * flux added as pga_diffflux adds it (alpha, trstate, hsp, grads, xcc and the face centres xf[a]), so
 * F[a] holds what pga_flux then pga_diffflux leave, to the bit. Replaces, for a stage, the reference's
 * three pg_flux launches (euler/kernels/flux/single.mako:4-9) and the three viscous ones
 * (navstokes/kernels/flux/alpha-damping.mako:3-99). */
int32_t pga_faces(const pga_case *c, pga_view Q, pga_view state, const pga_view *S, const pga_view *F,
                  const pga_box *boxes, int32_t viscous, double alpha, pga_view trstate, pga_view hsp, pga_view grads,
                  pga_view xcc, const pga_view *xf, void *stream);

/* dQ (= or +=) (iF_low + jF_low + kF_low - iF_high - jF_high - kF_high) * detJinv over the cells in box.
 * Replaces pg_applyflux: peregrine/simulator/advec/kernels/applyflux.mako:2-11. */
int32_t pga_applyflux(int32_t ne, pga_view dQ, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv,
                      int32_t append, pga_box box, void *stream);

/* the coupled chemistry source: dQ[0..4] = 0, dQ[5+n] = omega_n (kg/m^3/s).
 * Replaces pg_source: peregrine/simulator/chemint/kernels/source.mako:3-10. */
int32_t pga_source(const pga_case *c, pga_view Q, pga_view state, pga_view dQ, pga_box box, void *stream);

/* net production rates at n points: rho[n], Y[ns][n] (all species), T[n] -> omega[ns][n].
 * The body of pg_source without the solution layout (finiterate.mako production). */
int32_t pga_production(const pga_case *c, int32_t n, const double *rho, const double *Y, const double *T,
                       double *omega, void *stream);

/* out = [wout out] + sum_m w_m X_m, the last term times *dt when step (dt a device scalar).
 * Replaces pg_combine: peregrine/integrators/kernels/combine.mako:1-13. */
int32_t pga_combine(int32_t ne, pga_view out, int32_t own, double wout, int32_t nterms, const pga_view *X,
                    const double *w, const double *dt, int32_t step, pga_box box, void *stream);

/* the halo of a block periodic with itself along axis: each halo slab `depth` deep takes the
 * interior planes next to the opposite interface; the other axes run over the interior only.
 * Replaces pg_place on a self-periodic interface pair:
 * peregrine/multiblock/collections/exchange/kernels/place.mako:11-37. */
int32_t pga_fill_periodic(pga_view A, int32_t ncomp, const int32_t *ncell, int32_t ng, int32_t depth,
                          int32_t axis, void *stream);

/* a box of `extent` cells of every component copied from src to dst (at dst_start): the halo
 * slab of a block interface from its partner's interior planes. orient (3 digits, or null for
 * 1 2 3) is the interface's orientation: digit m names the source axis the destination's axis m
 * runs along, 1 2 3 for +i +j +k and 4 5 6 for -i -j -k; the cell dst_start + (a0, a1, a2) takes
 * the source cell src_start + sum_m a_m (+-1 along that axis). rot (9 doubles, row-major, or
 * null) is a rotational periodic's turn: the components [vec, vec + 3) are turned by it; with
 * rows = 1 the components are rows of three (a gradient each), every row is turned, and the rows
 * [vec, vec + 3) (a velocity gradient) are turned down each column too (vec -1: none). An exact
 * identity rot copies as null does. Replaces pg_place: place.mako:3-37 with
 * peregrine/multiblock/collections/exchange/collection.py:55-96 (haloview, nbrview, remap_across)
 * and peregrine/multiblock/interface/base.py:37-45,95-136 (nbrplane, signedaxis, nbrnface). */
int32_t pga_copy_box(pga_view dst, const int32_t *dst_start, pga_view src, const int32_t *src_start,
                     const int32_t *extent, int32_t ncomp, const int32_t *orient, const double *rot, int32_t vec,
                     int32_t rows, void *stream);

/* pga_copy_box for up to 6 slabs in one launch: slab n copies ncomp components of extent[3n..3n+2]
 * cells from src[n] (placed by src_start and the orientation digits orient[3n..], null: aligned) to
 * dst[n] at dst_start[3n..], turned by rot[9n..9n+8] (null, or an exact identity: a copy) as
 * pga_copy_box turns with vec and rows. The one-plane exchange of the gradients across every
 * interface of a block in one launch. Replaces, for those slabs, the reference's pg_place launches
 * (place.mako:3-37) that pga_copy_box replaces one by one. */
int32_t pga_copy_slabs(int32_t ntask, const pga_view *dst, const pga_view *src, const int32_t *dst_start,
                       const int32_t *src_start, const int32_t *extent, int32_t ncomp, const int32_t *orient,
                       const double *rot, int32_t vec, int32_t rows, void *stream);

/* 1 when the fused stage is compiled for the case's species count, else 0 (use the per-kernel
 * entry points) */
int32_t pga_stage_supported(const pga_case *c);

/* One RK stage of a block's interior in one pass after its three flux sweeps: the divergence of
 * the six face fluxes (iF, jF, kF from pga_flux on Qin), the chemistry's source (a reacting
 * case's rates, ns - 1 components per cell, from pga_rates or pga_substep with begin_dQ = 0),
 * the RK combine and the EOS of the result,
 * its Newton started from stin's T. Q0 is the step's start. The combine is
 * out = sum_t w[t] x_t over nterms terms in order, x_t = Qin (src 0), Q0 (1) or the stage's
 * dQ (2), a term with withdt[t] taking w[t] * (*dt) * x_t; dQ is never stored. Each cell reads
 * only itself, so Qout may be Qin and stout may be stin. kchunk: the k-planes one thread walks.
 * Replaces, for one stage, the reference's launches of pg_source
 * (chemint/kernels/source.mako:3-10), pg_applyflux (advec/kernels/applyflux.mako:2-11),
 * pg_combine (integrators/kernels/combine.mako:1-13) and the interior pg_fromQ
 * (euler/kernels/state/fromQ.mako:5-19), as sequenced by multiblock/solver.py:246-284. */
int32_t pga_stage(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view Qout, pga_view stout,
                  pga_view iF, pga_view jF, pga_view kF, pga_view detJinv, pga_view rates, int32_t nterms,
                  const int32_t *src, const double *w, const int32_t *withdt, const double *dt, pga_box box,
                  int32_t kchunk, void *stream);

/* A stage with two combines per cell (rk4's): pga_stage's divergence, source and EOS, with a first
 * combine into QA before the one into Qout. The first's terms read Qin (0), Q0 (1), dQ (2) or QA as it
 * stood (3); the second's read Qin, Q0, dQ or the value the first just took (4). Replaces, for one rk4
 * stage, pg_applyflux and the two pg_combine launches of peregrine/integrators/rungekutta.py:84-95 (the
 * running total, then the next stage or the step's end) and the interior pg_fromQ. */
int32_t pga_stage2(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view QA, pga_view Qout,
                   pga_view stout, pga_view iF, pga_view jF, pga_view kF, pga_view detJinv, pga_view rates,
                   int32_t nA, const int32_t *srcA, const double *wA, const int32_t *withdtA, int32_t nterms,
                   const int32_t *src, const double *w, const int32_t *withdt, const double *dt, pga_box box,
                   int32_t kchunk, void *stream);

/* pga_stage with its k-face fluxes computed in the same launch (the march): each thread walks a
 * chunk of kchunk cells of one (i, j) column, takes each k face's advective flux once (pga_flux's
 * arithmetic, from Qin, stin and the face areas kS) and keeps it for the two cells it lies between;
 * the i and j faces come from iF and jF (pga_flux along i and j first). No k-face array is written
 * or read. Results are pga_stage's after pga_flux along k, to the bit. The output must not be the
 * input (a chunk's last face reads cells the next chunk writes): PGA_EINVAL. For a block without
 * walls or viscous terms (their hooks change the k faces' fluxes). Replaces, for one stage, the
 * reference's k-axis pg_flux launch (euler/kernels/flux/single.mako:4-9) with the launches
 * pga_stage replaces. */
int32_t pga_stage_march(const pga_case *c, pga_view Qin, pga_view stin, pga_view Q0, pga_view Qout, pga_view stout,
                        pga_view iF, pga_view jF, pga_view kS, pga_view detJinv, pga_view rates, int32_t nterms,
                        const int32_t *src, const double *w, const int32_t *withdt, const double *dt, pga_box box,
                        int32_t kchunk, void *stream);

/* The carried species' production rates (kg/m^3/s) of every cell in box into rates (ns - 1
 * components): the coupled source of a reacting stage. The body of pg_source
 * (peregrine/simulator/chemint/kernels/source.mako:3-10) writing only the species rows. */
int32_t pga_rates(const pga_case *c, pga_view Q, pga_view state, pga_view rates, pga_box box, void *stream);

/* The halo pass of a stage: for up to 6 slabs, Q copied from src to dst (at dst_start), extent
 * cells each, the source cells placed by src_start and the slab's 3 orientation digits in orient
 * (null: all aligned) as in pga_copy_box, the momentum turned by the slab's 9 doubles in rot
 * (null, or an exact identity: not turned), then the EOS over the slab into stdst, its Newton
 * started from stguess's T at the same cells. Replaces the prerhs slot's pg_place and the halo
 * pg_fromQ (simulator/advec/simulator.py:57-64). */
int32_t pga_halo(const pga_case *c, int32_t ntask, const pga_view *Qdst, const pga_view *Qsrc,
                 const int32_t *dst_start, const int32_t *src_start, const int32_t *extent, const int32_t *orient,
                 const double *rot, pga_view stdst, pga_view stguess, void *stream);

/* --- the viscous terms (Navier-Stokes) ---------------------------------------------------- */

/* Per cell of up to 8 boxes: trstate = [mu, kappa, D_0 .. D_{ns-1}] by the case's transport model
 * (constant species values or kinetic-theory fits, Wilke's rule, the conductivity mean, Lewis
 * diffusivities) and hsp = the species enthalpies h_n(T) (J/kg).
 * Replaces pg_transport: peregrine/simulator/navstokes/kernels/transport.mako:5-35. */
int32_t pga_transport(const pga_case *c, pga_view Q, pga_view state, pga_view trstate, pga_view hsp,
                      const pga_box *boxes, int32_t nboxes, void *stream);

/* The viscous sponge after the transport, per cell of up to 8 boxes: trstate[0] (the viscosity)
 * times 1 + min(max(0, s (m - 1)), m - 1), s = (xcc . n - start) / length the cell centre's way
 * along the line, line = (nx, ny, nz, start, length, m) (host memory).
 * Replaces the viscsponge plugin's kernel: peregrine/plugins/simulator/kernels/viscsponge.mako:3-9,
 * peregrine/plugins/simulator/viscsponge.py:9-26. */
int32_t pga_viscsponge(pga_view xcc, pga_view trstate, const double *line, const pga_box *boxes, int32_t nboxes,
                       void *stream);

/* Per cell of box: grads[3 l + e] = d(row l)/dx_e for u, v, w, T and the carried Y, by central
 * differences in index space mapped by Jinv[3 r + e] = d xi_r / d x_e (one halo layer read).
 * Replaces pg_dqdxyz: peregrine/simulator/navstokes/kernels/dqdxyz.mako:16-41. */
int32_t pga_grads(const pga_case *c, pga_view Q, pga_view state, pga_view Jinv, pga_view grads, pga_box box,
                  void *stream);

/* pga_transport over its boxes and pga_grads over gbox in one launch, the same values as the two
 * calls (neither reads what the other writes; both read the halo pass's Q and state). For a step
 * of small blocks, where a launch costs more than its work. An empty gbox computes no gradients.
 * No counterpart of its own in the reference: the two kernels it replaces are
 * peregrine/simulator/navstokes/kernels/transport.mako:5-35 and dqdxyz.mako:16-41, launched one
 * after the other by peregrine/simulator/navstokes/simulator.py. */
int32_t pga_transport_grads(const pga_case *c, pga_view Q, pga_view state, pga_view trstate, pga_view hsp,
                            const pga_box *boxes, int32_t nboxes, pga_view Jinv, pga_view grads, pga_box gbox,
                            void *stream);

/* The alpha-damped viscous flux of the faces of one axis in box (face indices), added to F:
 * Newtonian stress, Fourier heat, Fick species fluxes with the correction velocity, their
 * enthalpy flux and the stress work; xcc the cell centres, xf the face centres.
 * Replaces pg_flux of navier-stokes diff-flux alpha-damping:
 * peregrine/simulator/navstokes/kernels/flux/alpha-damping.mako:3-99. */
int32_t pga_diffflux(const pga_case *c, int32_t axis, double alpha, pga_view Q, pga_view state, pga_view trstate,
                     pga_view hsp, pga_view grads, pga_view xcc, pga_view xf, pga_view S, pga_view F, pga_box box,
                     void *stream);

/* --- boundaries -------------------------------------------------------------------------- */

/* A wall: its kind and values (peregrine/simulator/euler/bcs.py:15-27, navstokes/bcs.py:27-53).
 * kind: 0 slp-adia-wall, 1 slp-isot-wall (Euler and Navier-Stokes); 2 no-slp-adia-wall,
 * 3 no-slp-isot-wall, 4 mov-adia-wall, 5 mov-isot-wall (Navier-Stokes). u: a moving wall's velocity;
 * T: an isothermal wall's temperature. A null pga_wall pointer is an adiabatic slip wall. */
typedef struct {
  int32_t kind;
  int32_t reserved;
  double u[3];
  double T;
} pga_wall;

/* A wall's halo on face `side` (0 low, 1 high) of `axis` of a block of ncell interior cells with ng
 * halo layers: each halo cell h planes out from the interior cell h planes in. A slip wall mirrors
 * the velocity's normal part (the wall face's unit normal from S); a no-slip or moving wall reflects
 * the velocity about the wall's, the energy carrying the change in its kinetic part; an adiabatic
 * wall copies the state, an isothermal one takes the image's p and Y at the wall's T by the EOS.
 * Replaces the preadvflux hooks: peregrine/simulator/euler/kernels/bcs/slp-{adia,isot}-wall.mako,
 * navstokes/kernels/bcs/{no-slp,mov}-{adia,isot}-wall.mako and wall.mako (stuckhalo, stuckhaloisot). */
int32_t pga_wall_mirror(const pga_case *c, int32_t axis, int32_t side, const pga_wall *w, const int32_t *ncell,
                        int32_t ng, pga_view Q, pga_view state, pga_view S, void *stream);

/* The flux through the same wall's faces, written over F: the pressure p_w = max(p + rho c u_n, 0)
 * of the cell next to the wall, and for a viscous case (trstate non-null) the stress of the two-point
 * normal velocity gradient relative to the wall (its normal part alone on a slip wall), the wall's
 * work and, on an isothermal wall, the heat flux from T_w with mu and kappa at the wall. Replaces
 * the postflux hooks: peregrine/simulator/euler/kernels/bcs/wall.mako,
 * navstokes/kernels/bcs/wall.mako (wallfaceflux, adiaflux, isotflux, wallcoeffs). */
int32_t pga_wall_flux(const pga_case *c, int32_t axis, int32_t side, const pga_wall *w, const int32_t *ncell,
                      int32_t ng, pga_view Q, pga_view state, pga_view trstate, pga_view xcc, pga_view xf, pga_view S,
                      pga_view F, void *stream);

/* The CFL reduction over the cells of `box`: out[0..2] (device memory) receive the largest
 * acoustic c sqrt(sum 1/d^2), convective sum |u . n| / d and combined (their sum) inverse time
 * scales, each folded by fmax from -inf; d the cell's lengths (dijk, infinite along an axis one
 * cell thick), n the mean of its two faces' unit normals per axis. `work` is device memory of at
 * least pga_cflmax_worksize(box) doubles for the per-workgroup partials. Replaces the cflmax
 * kernel and its reduction: peregrine/simulator/euler/kernels/cflmax.mako:3-33,
 * peregrine/multiblock/solver.py:292-296, peregrine/backend/device/backend.py:24-41. */
int64_t pga_cflmax_worksize(pga_box box);
int32_t pga_cflmax(const pga_case *c, pga_view Q, pga_view state, pga_view dijk, pga_view iS, pga_view jS,
                   pga_view kS, pga_box box, double *work, int64_t nwork, double *out, void *stream);

/* --- graphs and the all-finite check ----------------------------------------------------- */

typedef struct pga_graph pga_graph;

/* Capture the launches issued on `stream` (not the null stream) from pga_graph_begin to pga_graph_end
 * into a graph, instantiated for replay by pga_graph_launch on any stream. The library's launches
 * record no memset, allocation or host synchronisation, so a step's launches capture whole.
 * Replaces peregrine/backend/device/graph.py:10-47 and peregrine/backend/hip/driver.py:159-176
 * (hipStreamBeginCapture, hipStreamEndCapture, hipGraphInstantiate, hipGraphLaunch). */
int32_t pga_graph_begin(void *stream);
int32_t pga_graph_end(void *stream, pga_graph **out);
int32_t pga_graph_launch(const pga_graph *g, void *stream);
int32_t pga_graph_destroy(pga_graph *g);

/* out[0] (device int32) = 1 when every one of the ncomp components of every cell of box is finite, else
 * 0; set on the device by two launches (no memset, so graph-capturable). Replaces the nancheck plugin's
 * kernel and reduction: peregrine/plugins/solver/kernels/allfinite.mako:3-19,
 * peregrine/plugins/solver/nancheck.py. */
int32_t pga_allfinite(pga_view A, int32_t ncomp, pga_box box, int32_t *out, void *stream);

/* pga_cflmax with no work argument and no allocation: the partials go to a 6 MiB array the library
 * holds on each GPU for the life of the process (its address looked up on the first call on that
 * device). No counterpart in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Boxes up to 2^26 cells
 * (262144 partials); a larger one is refused with PGA_ERANGE. Calls on one GPU share the array, so
 * calls that overlap on one GPU (different streams, or threads) race; the array keeps the last
 * call's partials, each call writing every partial it reads. */
int32_t pga_cflmax_noalloc(const pga_case *c, pga_view Q, pga_view state, pga_view dijk, pga_view iS, pga_view jS,
                           pga_view kS, pga_box box, double *out, void *stream);

/* Sub-stepped chemistry's options: chem-integration.max-substeps, .bulk-substeps, .stiff-bins and
 * .equilibrium-tolerance (peregrine/simulator/chemint/substepped.py:14-33) */
typedef struct {
  int32_t max_substeps;  /* the cap on a cell's substeps in a step, >= 1 */
  int32_t bulk_substeps; /* the bulk pass's limit, 1..max_substeps; below the cap the stiff cells are listed */
  int32_t stiff_bins;    /* 1..16 bins of the stiff list, by the substep count of the previous call */
  int32_t reserved;
  double equilibrium_tolerance; /* 0 <= tol < 1; 0 for no equilibrium exit */
} pga_substep_opts;

/* Int32 entries of the stiff list a compacted call needs for `box`: a header of 1 + bins counts and
 * a segment per bin as long as the box; -1 for bad options or box. */
int64_t pga_substep_listsize(const pga_substep_opts *o, pga_box box);

/* Sub-stepped chemistry over the cells of `box` for a step of *dt (device): each cell integrates its
 * finite-rate chemistry in explicit substeps, each as long as its limiting species allows, with the
 * equilibrium exit at the first substep, and writes the mean rate over the step: into dQ (rows 0..4
 * zero, 5.. the rates) when begin_dQ, else into the ns - 1 rows of `out` as pga_rates does. The
 * substep counts are kept per cell in `substeps` (int32, device, box-shaped, i fastest) and key the
 * next call's bins; `list` (device int32, pga_substep_listsize entries) is needed when
 * bulk_substeps < max_substeps. Replaces chemint/substepped.py:52-68 (begin_dQ_graph) and the
 * kernels peregrine/simulator/chemint/kernels/{substep,bulk,stiff,reset}.mako. */
int32_t pga_substep(const pga_case *c, const pga_substep_opts *o, pga_view Q, pga_view state, pga_view out,
                    int32_t begin_dQ, int32_t *substeps, int64_t nsubsteps, int32_t *list, int64_t nlist,
                    const double *dt, pga_box box, void *stream);

/* Implicit chemistry's options: chem-integration.implicit-parts, .max-sweeps and
 * .refine-tolerance. No counterpart in the reference, whose chemistry is coupled or sub-stepped
 * (peregrine/simulator/chemint/{coupled,substepped}.py); this is the stiff alternative to the latter. */
typedef struct {
  int32_t parts;      /* linearly implicit Euler steps per step, >= 1 */
  int32_t max_sweeps; /* refinement corrections per linear solve, >= 1 */
  double tolerance;   /* a solve is done when every row's residual is within this share of the row's scale */
} pga_implicit_opts;

/* Doubles of device work pga_implicit needs for `box` (the FP64 matrices of 16-cell groups, the FP64
 * unknowns and two right-hand sides); -1 for a bad box or more than 11 species. */
int64_t pga_implicit_worksize(const pga_case *c, pga_box box);

/* Implicit chemistry over the cells of `box` for a step of *dt (device): each cell integrates the
 * same 0D system the sub-stepped form does (its ns mass fractions and T at constant density and
 * internal energy, chemint/kernels/substep.mako:14-50) by the linearly implicit Euler method in
 * o->parts parts, y1 = y0 + h x with (I - h J) x = f(y0), J by forward differences in FP64. Each
 * matrix is factored in FP32 on the matrix cores, 16 cells per wave, and in the same registers
 * each solve is refined against the FP64 matrix with FP64 residuals until every row's residual is
 * within o->tolerance of its scale |f_i| + sum_j |A_ij x_j|. It writes the mean rate over the step, rho (Y_end - Y_start) / dt,
 * into dQ (rows 0..4 zero) when begin_dQ, else into the ns - 1 rows of `out` as pga_rates does, and
 * each cell's refinement corrections over its parts into `sweeps` (int32, box-shaped, i fastest;
 * 1000 more for each solve left outside the tolerance). `work`: pga_implicit_worksize doubles of
 * device memory. Takes the place of pga_substep (the reference's chemint/substepped.py:52-68 slot)
 * for chem-integration.kind = implicit. */
int32_t pga_implicit(const pga_case *c, const pga_implicit_opts *o, pga_view Q, pga_view state, pga_view out,
                     int32_t begin_dQ, int32_t *sweeps, int64_t nsweeps, double *work, int64_t nwork,
                     const double *dt, pga_box box, void *stream);

/* Doubles of device work pga_lu_solve needs for n systems of nv unknowns; -1 when nv is not 1..12. */
int64_t pga_lu_worksize(int64_t n, int32_t nv);

/* n dense systems A x = b (device memory; A row-major [n][nv][nv], b and x [n][nv], FP64), nv <= 12:
 * each A rounded to FP32 and factored without pivoting by the batched LU on the matrix cores, then
 * x refined by x += (LU)^-1 (b - A x) with the residual in FP64 until every row's residual is within
 * `tolerance` of |b_i| + sum_j |A_ij x_j| or max_sweeps corrections are taken; sweeps (or null)
 * receives each system's corrections (max_sweeps + 1 when still outside the tolerance). The linear
 * kernel of pga_implicit, exposed on its own. No counterpart in the reference. */
int32_t pga_lu_solve(int64_t n, int32_t nv, const double *A, const double *b, double *x, int32_t max_sweeps,
                     double tolerance, int32_t *sweeps, double *work, int64_t nwork, void *stream);

/* --- launches over a table of blocks ------------------------------------------------------ */

/* One block of a batch: its registers (the first nregs of Q), state, face areas and face flux
 * arrays per axis, 1/J and (reacting) rates, its interior box, and its halo slabs as pga_halo takes
 * them: slab t fills slab_extent[t] cells at slab_dst[t] of this block from block slab_src[t] of the
 * batch, read from slab_from[t] with orientation digits slab_orient[t] (1 2 3 aligned). Tile shapes
 * (0 for the defaults): stage_wi the stage's tile width along i (16, 32 or 64; its tiles are
 * stage_wi x 256 / stage_wi columns), stage_kc the k-planes a stage thread walks, flux_ki the k-sweep's
 * tile width along i (16, 32 or 64; 256 / flux_ki faces along k). */
typedef struct {
  pga_view Q[4];
  pga_view state, iS, jS, kS, iF, jF, kF, detJinv, rates;
  pga_box interior;
  int32_t nslab;
  int32_t slab_src[6];
  int32_t slab_dst[6][3], slab_from[6][3], slab_extent[6][3], slab_orient[6][3];
  int32_t stage_wi, stage_kc, flux_ki, reserved;
} pga_batch_block;

typedef struct pga_batch pga_batch;

/* Bytes of the device table a batch of nblocks blocks with nslabs halo slabs in all needs. */
int64_t pga_batch_bytes(int32_t nblocks, int32_t nslabs);

/* A batch over `table` (device memory of pga_batch_bytes, the caller's, uploaded here once): checks
 * every block as the block-at-a-time entry points check theirs, and writes a host handle to *out.
 * The arrays the blocks name must outlive the batch; the library frees nothing of the caller's.
 * Replaces the reference's per-kernel range tables over a rank's blocks (peregrine/launch/table.py:
 * 83-106, policy.mako:8-13) for the fused stage's phases. */
int32_t pga_batch_create(const pga_case *c, int32_t nblocks, const pga_batch_block *blocks, int32_t nregs,
                         void *table, int64_t nbytes, pga_batch **out);
int32_t pga_batch_destroy(pga_batch *b);

/* The fused stage's phases over every block in one launch each, each block's result the
 * block-at-a-time entry point's to the bit: the halo pass of register q (pga_halo without turns),
 * the coupled rates of register q into each block's rates (pga_rates), the flux sweep of one axis
 * (pga_flux), and the stage from register q into register qout with the step's start in register 0
 * (pga_stage: divergence, source, combine, then the EOS into each block's state). */
int32_t pga_batch_halo(const pga_batch *b, int32_t q, void *stream);
int32_t pga_batch_rates(const pga_batch *b, int32_t q, void *stream);
int32_t pga_batch_flux(const pga_batch *b, int32_t q, int32_t axis, void *stream);
int32_t pga_batch_stage(const pga_batch *b, int32_t q, int32_t qout, int32_t nterms, const int32_t *src,
                        const double *w, const int32_t *withdt, const double *dt, void *stream);

/* Implicit chemistry (pga_implicit's method and kernels) over every block's cells of register q in one
 * launch per phase, the mean rates into each block's rates array; sweeps (int32) and work (doubles,
 * pga_batch_implicit_worksize) are device memory laid out by the batch's cell tiles. No counterpart
 * in the reference. */
int64_t pga_batch_implicit_worksize(const pga_batch *b);
int32_t pga_batch_implicit(const pga_batch *b, const pga_implicit_opts *o, int32_t q, int32_t *sweeps,
                           int64_t nsweeps, double *work, int64_t nwork, const double *dt, void *stream);

#ifdef __cplusplus
}
#endif

#endif
