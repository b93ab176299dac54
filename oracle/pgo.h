/*
 * pgo: the CPU oracle of the per-stage right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain C restatement of the reference's
 * hot-path kernels, written loop by loop from the Mako templates they are
 * rendered from (each function cites its template). It is read by the tests,
 * by the smoke check and by the CPU-baseline leg of the benchmark, always as
 * the checker; the product path never links, loads or calls it.
 *
 * Layout: every array is the device layout of the reference, Fortran order
 * [i, j, k, comp] with i fastest and each component a separate plane, ng halo
 * layers on every side; a view carries the j, k and component strides.
 */
#ifndef PGO_H
#define PGO_H

#include <stdint.h>

typedef struct {
  double *d;
  int64_t sj, sk, sc;
} pgo_arr;

typedef struct {
  int start[3];
  int extent[3];
} pgo_box;

/* what the kernels read from the mixture, and which kernel variants a case takes */
typedef struct {
  int ns, ne;
  int eos;     /* 0 cpg, 1 tpg */
  int recon;   /* 0 constant, 1 muscl */
  int limiter; /* 1 koren 2 minmod 3 mc 4 superbee 5 vanleer */
  int primary; /* 0 kepaec, 1 hllc, 2 rusanov */
  int nt, ng_; /* cppoly terms (7), gpoly terms (8) */
  const double *MW, *MWinv, *Rsp, *cppolyR, *h298R, *cp0, *gpoly, *h298;
  /* finite-rate tables, one row per reaction (width 4 for species lists) */
  int nr, ntb;
  const int32_t *kind, *reversible, *product, *tb, *troeidx;
  const int32_t *nfwd, *fwdsp, *nreact, *reactsp, *nprod, *prodsp, *nnet, *netsp, *ntbeff, *tbeffsp;
  const double *logA, *b, *EaR, *logprA, *prb, *prEaR, *nusum;
  const double *fwdorder, *reactnu, *prodnu, *netnu, *netnuMW, *lfcpoly, *tbdefeff, *tbeffdev;
  double logprefRu, prefRu, Rupref, logcfloor, expclamp, troeuc, troeuhinv, cfloor;
  /* transport: -1 none (Euler), 0 constant, 1 kinetic theory, 2 Chung (dense, on kinetic
   * theory); Wilke mixing, Lewis diffusion */
  int trans, nmu, nkappa;
  const double *hpoly, *mu0, *kappa0, *lewis, *MWqinv, *wilkepair, *mupoly, *kappapoly;
  double Ru;
  /* the cubic (eos 2): u, w, sigma, eps of v^2 + u b v + w b^2, and per species constants */
  double cubic[4];
  const double *cubicsqrtac, *cubicb, *cubicda, *cubicmTc, *cubiconepm, *cubicsqrtacMW, *cubicbMW, *vshift,
      *vshiftMW, *cubicdaMW, *acentric;
  /* Chung: cp/R fits, sigma (angstrom), epsilon/k (K), the E (10 x 4) and B (7 x 4) tables */
  const double *cppoly, *chungsigma, *chungepsk, *chungvisc, *chungcond;
  /* binary diffusion (diffusion 1; 0 is Lewis numbers): per species n and partner slot k < width,
   * pairfit[(n ns + k) npair + t] (T^3/2 / D at unit pressure in ln T) and bathof[n ns + k] the
   * partner, selffit[n npair + t] its own pair; partial: a bath short of every species */
  int diffusion, npair, diffwidth, diffpartial;
  const double *pairfit, *selffit, *bathof;
} pgo_case;

double pgo_tlog(double x);
double pgo_elog(double x);
double pgo_texp(double x);
double pgo_dcbrt(double x);

void pgo_fromQ(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_box r);
void pgo_fromprims(const pgo_case *c, pgo_arr prims, pgo_arr Q, pgo_arr state, pgo_box r);
void pgo_flux(const pgo_case *c, int axis, pgo_arr Q, pgo_arr state, pgo_arr S, pgo_arr F, pgo_box r);
/* recon/{koren,minmod,mc,superbee,vanleer}.mako: b phi(a / b) for b > 0, kind 1..5 as the case's limiter */
double pgo_limit(int kind, double a, double b);
/* a face's two reconstructed sides (right cell (i, j, k) along axis), each p, T, rho, e, c, u, v, w, Y */
void pgo_sides(const pgo_case *c, pgo_arr Q, pgo_arr st, int axis, int i, int j, int k, double *out);
void pgo_applyflux(int ne, pgo_arr dQ, pgo_arr iF, pgo_arr jF, pgo_arr kF, pgo_arr detJinv, int append, pgo_box r);
void pgo_production(const pgo_case *c, double rho, const double *Y, double T, double *omega);
/* omega and the destruction rates D (kg/m^3/s) in one fold; D may be null */
void pgo_sources(const pgo_case *c, double rho, const double *Y, double T, double *omega, double *D);
double pgo_dlog1p(double x);
/* sub-stepped chemistry, one pass per cell (chemint/kernels/substep.mako, bulk.mako): the mean
 * rate over dt into dQ (begin_dQ) or the ns - 1 rows of out, each cell's count into substeps */
void pgo_substep(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr out, int begin_dQ, int32_t *substeps, int cap,
                 double eqtol, double dt, pgo_box r);
void pgo_source(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr dQ, pgo_box r);
/* implicit chemistry (the FP64 plain form): the linearly implicit Euler method in `parts` parts
 * over dt per cell, its systems solved by an FP64 LU without pivoting; the mean rate into dQ
 * (begin_dQ) or the ns - 1 rows of out */
void pgo_implicit(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr out, int begin_dQ, int parts, double dt,
                  pgo_box r);
/* its pieces, for the tests: one cell advanced in place, the 0D system's f, the matrix I - h J, the
 * forward-difference step, and the unpivoted LU and its solve */
void pgo_implicit_cell(const pgo_case *c, double rho, double *y, double dt, int parts);
void pgo_implicit_rhs(const pgo_case *c, double rho, const double *y, double *f);
void pgo_implicit_matrix(const pgo_case *c, double rho, const double *y, double h, double *A);
double pgo_fd_step(double ym, int temperature);
void pgo_lu_nopivot(int n, double *A);
void pgo_lu_apply(int n, const double *LU, double *z);

void pgo_combine(int ne, pgo_arr out, int own, double wout, int nterms, const pgo_arr *X, const double *w,
                 double dt, int step, pgo_box r);
void pgo_fill_periodic(pgo_arr A, int ncomp, const int *ncell, int ng, int depth, int axis);

/* the viscous terms: transport state (mu, kappa, D_n) and species enthalpies per cell; the
 * cell gradients of u, v, w, T and the carried Y; the alpha-damped viscous face flux, added to F */
void pgo_transport(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr trstate, pgo_arr hsp, pgo_box r);
void pgo_grads(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr Jinv, pgo_arr grads, pgo_box r);
void pgo_diffflux(const pgo_case *c, int axis, double alpha, pgo_arr Q, pgo_arr state, pgo_arr trstate,
                  pgo_arr hsp, pgo_arr grads, pgo_arr xcc, pgo_arr xf, pgo_arr S, pgo_arr F, pgo_box r);

/* one whole step of a single periodic block: the stage sequence of the reference's solver */
typedef struct {
  int ncell[3], ng;
  pgo_arr regs[4]; /* solution registers; the last is dQ */
  int nregs;
  pgo_arr state, iF, jF, kF, iS, jS, kS, detJinv;
  int integrator; /* 0 rk1, 1 rk2, 2 tvd-rk3, 3 rk4 */
  int chemistry;  /* 1: the coupled source begins dQ; 2: sub-stepped chemistry does; 3: implicit chemistry does */
  /* sub-stepped chemistry: the cap, the equilibrium tolerance, the counts per interior cell, and the
   * step's dt (set by pgo_stage; pgo_rhs reads the last one) */
  int subcap;
  double subeqtol;
  int32_t *substeps;
  double dt;
  /* Navier-Stokes: the viscous arrays and the damping coefficient */
  int viscous;
  double alpha;
  pgo_arr trstate, hsp, grads, Jinv, xcc, xf[3];
  /* per face (i low, i high, j low, ...): 0 periodic, else 1 + the wall's kind (0 slp-adia,
   * 1 slp-isot, 2 no-slp-adia, 3 no-slp-isot, 4 mov-adia, 5 mov-isot) with its velocity and T;
   * nstate the state fields (9 for the real gas) */
  int bc[6];
  int nstate;
  double bcwall[6][3], bcT[6];
  /* the viscous sponge (plugins/simulator/viscsponge.py): on when sponge, its line as
   * (nx, ny, nz, start, length, multiplier) */
  int sponge;
  double spongeline[6];
  /* implicit chemistry (chemistry 3): linearly implicit Euler parts per step */
  int imparts;
} pgo_block;

/* plugins/simulator/kernels/viscsponge.mako: the viscosity trstate[0] raised along the line
 * (nx, ny, nz, start, length, multiplier): times 1 + min(max(0, s (m - 1)), m - 1), s the cell
 * centre's distance along the normal from start over length */
void pgo_viscsponge(pgo_arr xcc, pgo_arr trstate, const double *line, pgo_box r);

/* the walls (simulator/euler/bcs.py:15-27, navstokes/bcs.py:27-53): the halo as the interior's
 * image across the wall face (side 0 the low face of axis, 1 the high), and the wall's own flux
 * through its faces; kind 0 slp-adia, 1 slp-isot, 2 no-slp-adia, 3 no-slp-isot, 4 mov-adia,
 * 5 mov-isot, wall the wall's velocity, Tw its temperature */
void pgo_wall_mirror(const pgo_case *c, int axis, int side, int kind, const double *wall, double Tw, const int *ncell,
                     int ng, int nstate, pgo_arr Q, pgo_arr state, pgo_arr S);
void pgo_wall_flux(const pgo_case *c, int axis, int side, int kind, const double *wall, double Tw, const int *ncell,
                   int ng, int viscous, pgo_arr Q, pgo_arr state, pgo_arr trstate, pgo_arr xcc, pgo_arr xf, pgo_arr S,
                   pgo_arr F);

/* the largest acoustic, convective and combined inverse time scales over the box's cells (the
 * CFL controller's step is cfl over the combined one) */
void pgo_cflmax(pgo_arr Q, pgo_arr state, pgo_arr dijk, pgo_arr iS, pgo_arr jS, pgo_arr kS, pgo_box r, double out[3]);

void pgo_step(const pgo_case *c, pgo_block *b, double dt);
/* one right-hand side of register q into register dq (prerhs then rhs), as a stage begins it */
void pgo_rhs(const pgo_case *c, pgo_block *b, int q, int dq);
void pgo_stage(const pgo_case *c, pgo_block *b, int n, double dt);
/* the RK combine(s) of stage n alone, over the interior */
void pgo_advance(const pgo_case *c, pgo_block *b, int n, double dt);
int pgo_threads(void);

#endif
