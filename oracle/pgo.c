/*
 * pgo: the CPU oracle of the per-stage right-hand side (see pgo.h).
 *
 * TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's
 * CPU-baseline leg call it as the checker; the product never does.
 *
 * Built with -ffp-contract=off so every expression is evaluated in the order
 * written, one rounding per operation, as the reference's host backend does
 * on a target without fused multiply-add.
 */
#include "pgo.h"

#define PGA_ELOG_TABLE static const
#include "pga_elog_tables.h"

#include <float.h>
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The specified methods (ln by elog, cube roots by dcbrt, ln(1 + x) by dlog1p, explicit fused
 * multiply-adds in every Horner step and mixture sum) that the device takes too; built with
 * -DPGO_LIBM instead (libpgo_libm.so), the oracle takes the reference's host calls, libm's log, cbrt
 * and log1p, and plain products (one rounding each, -ffp-contract=off), so the gap between the two
 * builds measures what the specified methods move */
#ifdef PGO_LIBM
#define ELOG(x) log(x)
#define DCBRT(x) cbrt(x)
#define DLOG1P(x) log1p(x)
#define FMA(a, b, c) ((a) * (b) + (c))
#else
#define ELOG(x) pgo_elog(x)
#define DCBRT(x) pgo_dcbrt(x)
#define DLOG1P(x) pgo_dlog1p(x)
#define FMA(a, b, c) fma(a, b, c)
#endif

#define MAXNS 16
#define MAXNR 64
#define MAXTB 16

static inline double *at(pgo_arr a, int i, int j, int k, int c) {
  return a.d + i + j * a.sj + k * a.sk + c * a.sc;
}

int pgo_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------
 * tlog / texp: backend/base/mathfns.py:1-22, backend/base/kernels/mathfns.mako:6-54
 * log by a 128-bin table over the mantissa and a 5-term series; exp by a 64-bin
 * table of 2^(j/64) and a 6-term series, the exponent added in the bits.
 * ---------------------------------------------------------------------- */
static double LOGTAB[128], RCPTAB[128], EXPTAB[64];
static int tables_ready = 0;
static const double LN2 = 0.6931471805599453;
/* 64/ln 2 and ln 2/64 in two parts, as correctly rounded from a 40-digit ln 2 */
static const double EXPK = 0x1.71547652b82fep+6;
static const double LN2BINS_HI = 0x1.62e42fef00000p-7;
static const double LN2BINS_LO = 0x1.473de6af278edp-40;

static void make_tables(void) {
  if (tables_ready) return;
  for (int k = 0; k < 128; k++) {
    double m = 0.5 + (k + 0.5) / (2 * 128);
    LOGTAB[k] = log(m);
    RCPTAB[k] = 1.0 / m;
  }
  for (int j = 0; j < 64; j++) EXPTAB[j] = pow(2.0, j / 64.0);
  tables_ready = 1;
}

double pgo_tlog(double x) {
  make_tables();
  int64_t bits;
  memcpy(&bits, &x, sizeof bits);
  const int expo = (int)((bits >> 52) & 0x7ff) - 1022;
  const int bin = (int)((bits >> (52 - 7)) & 127);
  bits = (bits & 0x000fffffffffffffLL) | 0x3fe0000000000000LL;
  double mant;
  memcpy(&mant, &bits, sizeof mant);
  const double u = mant * RCPTAB[bin] - 1.0;
  double s = 0.2;
  s = s * u - 0.25;
  s = s * u + 1.0 / 3.0;
  s = s * u - 0.5;
  s = s * u + 1.0;
  return expo * LN2 + LOGTAB[bin] + u * s;
}

double pgo_texp(double x) {
  make_tables();
  const double t = x * EXPK + 6755399441055744.0;
  int64_t bits;
  memcpy(&bits, &t, sizeof bits);
  const int kexp = (int)bits, nexp = kexp >> 6;
  const double kd = t - 6755399441055744.0;
  const double r = (x - kd * LN2BINS_HI) - kd * LN2BINS_LO;
  double s = 1.0 / 120.0;
  s = s * r + 1.0 / 24.0;
  s = s * r + 1.0 / 6.0;
  s = s * r + 0.5;
  s = s * r + 1.0;
  s = s * r + 1.0;
  const double m = EXPTAB[kexp & 63] * s;
  memcpy(&bits, &m, sizeof bits);
  bits += (int64_t)(nexp > -1022 ? nexp : -1022) << 52;
  double y;
  memcpy(&y, &bits, sizeof y);
  return y;
}

/* ------------------------------------------------------------------------
 * ln T of the equation of state. The reference takes it from the platform's log
 * (caloric/tdepcp/kernels/caloric.mako:5); libm logs differ in the last bit between
 * platforms, and the degree-6 ln T polynomials turn a last-bit change into 1e-11 of h.
 * This build fixes the method on every platform: x = m 2^e, m in [1, 2); r_k a
 * 9-bit reciprocal of m's bin centre (top 7 bits); ln x = e ln 2 - ln r_k + ln(1 + r),
 * r = m r_k - 1 by one fma, ln(1 + r) by its series to r^7; -ln r_k and ln 2 in two
 * parts. Tables: pga_elog_tables.h, written once by the build.
 * ---------------------------------------------------------------------- */
double pgo_elog(double x) {
  /* zero, negative, NaN and infinity: the limits; a subnormal is scaled into the normal range */
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / 0.0;
  if (x > DBL_MAX) return x;
  int escale = 0;
  if (x < DBL_MIN) x *= 18014398509481984.0, escale = -54; /* 2^54 */
  if (escale == 0 && fabs(x - 1.0) < 0.03125) {
    /* near 1 the table's terms would cancel: ln(1 + r) by its series, r = x - 1 exact */
    const double r = x - 1.0;
    double q = -1.0 / 14.0;
    for (int n = 13; n >= 2; n--) q = fma(q, r, ((n & 1) ? 1.0 : -1.0) / n);
    return r + (r * r) * q;
  }
  int64_t bits;
  memcpy(&bits, &x, sizeof bits);
  const int e = (int)(bits >> 52) - 1023 + escale;
  const int k = (int)((bits >> 45) & 127);
  int64_t mb = (bits & 0x000fffffffffffffLL) | 0x3ff0000000000000LL;
  double m;
  memcpy(&m, &mb, sizeof m);
  const double r = fma(m, ELOG_RCP[k], -1.0);
  double q = 1.0 / 7.0;
  q = fma(q, r, -1.0 / 6.0);
  q = fma(q, r, 0.2);
  q = fma(q, r, -0.25);
  q = fma(q, r, 1.0 / 3.0);
  q = fma(q, r, -0.5);
  const double p = (r * r) * q;
  const double hi = e * PGA_ELOG_LN2HI + ELOG_LHI[k];
  const double lo = e * PGA_ELOG_LN2LO + ELOG_LLO[k];
  return hi + (r + (p + lo));
}

/* ------------------------------------------------------------------------
 * the equation of state
 * ---------------------------------------------------------------------- */
typedef struct {
  double p, T, rhoinv, rho, e, cp, gamma, c, Z;
} eosstate;

typedef struct {
  double R, cp, h298, cpcoeff[8], b, c, da;
} mixmom;

/* eos/kernels/mixing.mako:3-30: the carried species, then the remainder. The sums, and the
 * Horner steps of the ln T polynomials below, are taken as explicit fused multiply-adds: the
 * polynomials' terms cancel by six orders, so their rounding sequence is fixed by the code (the
 * device evaluates the same sequence) rather than left to a compiler's contraction. */
static void mix(const pgo_case *c, const double *Y, mixmom *m) {
  memset(m, 0, sizeof *m);
  double Ylast = 1.0;
  for (int n = 0; n < c->ns; n++) {
    double Yn;
    if (n < c->ns - 1) {
      Yn = Y[n];
      Ylast -= Yn;
    } else {
      Yn = Ylast;
    }
    m->R = FMA(Yn, c->Rsp[n], m->R);
    if (c->eos == 0) {
      m->cp = FMA(Yn, c->cp0[n], m->cp);
    } else {
      for (int t = 0; t < c->nt; t++) m->cpcoeff[t] = FMA(Yn, c->cppolyR[n * c->nt + t], m->cpcoeff[t]);
    }
    m->h298 = FMA(Yn, c->h298R[n], m->h298);
    if (c->eos == 2) {
      m->b = FMA(Yn, c->cubicbMW[n], m->b);
      m->c = FMA(Yn, c->vshiftMW[n], m->c);
      m->da = FMA(Yn, c->cubicdaMW[n], m->da);
    }
  }
}

/* caloric/tdepcp/kernels/caloric.mako:4-15 */
static void ideal_cph(const pgo_case *c, const mixmom *m, double T, double *cp, double *h) {
  const double u = ELOG(T);
  double cpv = 0.0, hv = 0.0, hc = 0.0;
  for (int t = c->nt - 1; t >= 0; t--) {
    cpv = FMA(cpv, u, m->cpcoeff[t]);
    hc = FMA(-(double)(t + 1), hc, m->cpcoeff[t]);
    hv = FMA(hv, u, hc);
  }
  *cp = cpv;
  *h = FMA(hv, T, m->h298);
}

/* ------------------------------------------------------------------------
 * the cubic real gas: eos/realgas/kernels/eos.mako, newton.mako. Logarithms are taken by the
 * EOS's log (pgo_elog), as everywhere in this path.
 * ---------------------------------------------------------------------- */
typedef struct {
  double MWmix, bm, cm, S1, S2, am, dam, Astar, Bstar, Cstar, invRuT, Z, z1, z2, invslope, dZdT, v, I, logratio, q, qd;
} cubicpt;

#define CU (c->cubic[0])
#define CW (c->cubic[1])
#define CSIG (c->cubic[2])
#define CEPS (c->cubic[3])

/* Soave's |1 + m (1 - sqrt(T / Tc))| as |(1 + m) - sqrt(T) m / sqrt(Tc)| */
static double modulus(const pgo_case *c, int n, double sqrtT) { return fabs(FMA(-sqrtT, c->cubicmTc[n], c->cubiconepm[n])); }

/* eos.mako:13-18 */
static void cubicmix(const pgo_case *c, const mixmom *m, cubicpt *q) {
  q->MWmix = c->Ru / m->R;
  q->bm = m->b * q->MWmix;
  q->cm = m->c * q->MWmix;
  q->S2 = m->da * q->MWmix;
}

/* eos.mako:21-32: k_ij = 0, am = S1^2, dam/dT = -S1 S2 / sqrt(T) */
static void attraction(const pgo_case *c, cubicpt *q, const double *Y, double T) {
  const double sqrtT = sqrt(T);
  double s1 = 0.0, Ylast = 1.0;
  for (int n = 0; n < c->ns - 1; n++) {
    const double Yn = Y[n];
    Ylast -= Yn;
    s1 = FMA(Yn, c->cubicsqrtacMW[n] * modulus(c, n, sqrtT), s1);
  }
  s1 = FMA(Ylast, c->cubicsqrtacMW[c->ns - 1] * modulus(c, c->ns - 1, sqrtT), s1);
  q->S1 = s1 * q->MWmix;
  q->am = q->S1 * q->S1;
  q->dam = -q->S1 * q->S2 / sqrtT;
}

/* eos.mako:34-43 */
static void dimless(const pgo_case *c, cubicpt *q, double p, double T) {
  q->invRuT = 1.0 / (c->Ru * T);
  const double pRT = p * q->invRuT;
  q->Astar = q->am * pRT * q->invRuT;
  q->Bstar = q->bm * pRT;
  q->Cstar = q->cm * pRT;
  const double B = q->Bstar, B2 = B * B;
  q->z1 = q->Astar + CW * B2 - CU * B - CU * B2;
  q->z2 = -(1.0 + B - CU * B);
}

/* eos.mako:86-89 */
static void rootslope(cubicpt *q) { q->invslope = 1.0 / (3.0 * q->Z * q->Z + 2.0 * q->Z * q->z2 + q->z1); }

/* the cube root by an exponent split and Newton steps in plain IEEE arithmetic: the device's
 * dcbrt, the same sequence (x = a 2^(3q), a in [0.5, 4), six steps from a linear start) */
double pgo_dcbrt(double x) {
  if (x == 0.0 || !(fabs(x) <= DBL_MAX)) return x;
  int e;
  const double m = frexp(fabs(x), &e);
  int q = e / 3, r = e - 3 * q;
  if (r < 0) r += 3, q -= 1;
  const double a = ldexp(m, r);
  double y = 0.75 + 0.25 * a;
  for (int k = 0; k < 6; k++) y = y - (y * y * y - a) / (3.0 * (y * y));
  return copysign(ldexp(y, q), x);
}

/* eos.mako:46-83: the cubic's root; of the roots above B*, the least Gibbs energy departure's.
 * One real root: Cardano (with pgo_dcbrt). Three: the largest by Newton from its upper bound
 * 2 sqrt(Q) - z2/3, the other two from the deflated quadratic -- the roots the reference takes by
 * the trigonometric form, here in IEEE operations only so the device gives the same bits. */
static void zsolve(const pgo_case *c, cubicpt *q, double p, double T) {
  dimless(c, q, p, T);
  const double B = q->Bstar, B2 = B * B, z1 = q->z1, z2 = q->z2;
  const double z0 = -(q->Astar * B + CW * B2 * (1.0 + B));
  const double Q = (z2 * z2 - 3.0 * z1) * (1.0 / 9.0);
  const double R = (2.0 * z2 * z2 * z2 - 9.0 * z2 * z1 + 27.0 * z0) * (1.0 / 54.0);
  const double M = R * R - Q * Q * Q;
  const double z2o3 = z2 * (1.0 / 3.0);
  if (M > 0.0) {
    const double S = -copysign(DCBRT(fabs(R) + sqrt(M)), R);
    q->Z = S + Q / S - z2o3;
  } else {
    double r0 = 2.0 * sqrt(Q) - z2o3;
    for (int k = 0; k < 100; k++) {
      const double f = ((r0 + z2) * r0 + z1) * r0 + z0;
      const double df = (3.0 * r0 + 2.0 * z2) * r0 + z1;
      if (!(df > 0.0)) break;
      const double dr = f / df;
      r0 -= dr;
      if (!(fabs(dr) > 1e-15 * fabs(r0))) break;
    }
    const double bq = z2 + r0, cq = z1 + bq * r0, disc = bq * bq - 4.0 * cq;
    double r1 = r0, r2 = r0;
    if (disc >= 0.0) {
      const double qq = -0.5 * (bq + copysign(sqrt(disc), bq));
      const double ra = qq, rb = qq != 0.0 ? cq / qq : 0.0;
      r1 = fmax(ra, rb);
      r2 = fmin(ra, rb);
    }
    const double roots[3] = {r0, r1, r2};
    double Z = r0;
    int above = 0;
    for (int k = 0; k < 3; k++) above += roots[k] > B;
    if (above > 1) {
      const double AoBse = q->Astar / (B * (CSIG - CEPS));
      double gmin = 0.0;
      int found = 0;
      for (int k = 0; k < 3; k++) {
        const double z = roots[k];
        if (z <= B) continue;
        const double g = z - 1.0 - ELOG(z - B) + AoBse * ELOG((z + CEPS * B) / (z + CSIG * B));
        if (!found || g < gmin) {
          gmin = g;
          Z = z;
          found = 1;
        }
      }
    }
    q->Z = Z;
  }
  rootslope(q);
}

/* eos.mako:91-97: dZ for given dA*, dB* */
static double dZ(const pgo_case *c, const cubicpt *q, double dA, double dB) {
  const double B = q->Bstar, Z = q->Z, Z2 = Z * Z;
  const double dz0 = -(B * dA + (q->Astar + (2.0 * B + 3.0 * B * B) * CW) * dB);
  const double dz1 = dA + (2.0 * B * (CW - CU) - CU) * dB;
  const double dz2 = -(1.0 - CU) * dB;
  return -(dz2 * Z2 + dz1 * Z + dz0) * q->invslope;
}

/* eos.mako:100-115 */
static void departures(const pgo_case *c, cubicpt *q, double T, double *hdep, double *cpdep) {
  const double Tinv = c->Ru * q->invRuT;
  const double dlnamdT = q->dam / q->am;
  q->dZdT = dZ(c, q, -2.0 * q->Astar * Tinv * (1.0 - 0.5 * T * dlnamdT), -q->Bstar * Tinv);
  const double x = q->Z / q->Bstar, xm1 = x - 1.0;
  const double D = x * x + CU * x + CW;
  q->logratio = ELOG((x + CEPS) / (x + CSIG));
  const double invBmRu = 1.0 / (q->bm * c->Ru);
  const double hres = invBmRu / (CSIG - CEPS) * (q->am * Tinv - q->dam) * q->logratio;
  q->qd = q->dam * invBmRu;
  q->q = q->am * invBmRu * Tinv;
  const double num = D - q->qd * xm1;
  *cpdep = hres * 0.5 * T * dlnamdT + num * num / (D * D - q->q * (2.0 * x + CU) * xm1 * xm1) - 1.0;
  *hdep = hres + (q->Z - 1.0) - q->Cstar;
}

/* eos.mako:117-131 */
static void gamma_c(const pgo_case *c, const cubicpt *q, double p, double T, double rho, double cp, double invZt,
                    double *gamma, double *cs) {
  const double invp = 1.0 / p, Tinv = c->Ru * q->invRuT;
  const double dZtdT = q->dZdT + q->Cstar * Tinv;
  const double dZdp = dZ(c, q, q->Astar * invp, q->Bstar * invp);
  const double dZtdp = dZdp - q->Cstar * invp;
  const double drhodp = rho * invp * (1.0 - p * dZtdp * invZt);
  const double r = -Tinv * (1.0 + T * dZtdT * invZt);
  const double csq = 1.0 / (drhodp - (T / cp) * r * r);
  *gamma = drhodp * csq;
  *cs = sqrt(fabs(csq));
}

/* cpg eos/cpg/kernels/eos.mako:5-15; tpg eos/tpg/kernels/eos.mako:8-20 and newton.mako:3-11;
 * real gas eos/realgas/kernels/eos.mako:157-189 with newton.mako:3-15 */
static void fromrhoE(const pgo_case *c, const mixmom *m, const double *Y, double rho, double e, double Tguess,
                     eosstate *s) {
  if (c->eos == 0) {
    const double cv = m->cp - m->R;
    s->rho = rho;
    s->rhoinv = 1.0 / rho;
    s->e = e;
    s->T = (e - m->h298) / cv;
    s->p = rho * m->R * s->T;
    s->cp = m->cp;
    s->gamma = m->cp / cv;
    s->c = sqrt(s->gamma * m->R * s->T);
    s->Z = 1.0;
    return;
  }
  const double tol = 1e5 * DBL_EPSILON;
  if (c->eos == 2) {
    cubicpt q;
    cubicmix(c, m, &q);
    q.v = q.MWmix / rho + q.cm;
    q.I = ELOG((q.v + CSIG * q.bm) / (q.v + CEPS * q.bm)) / ((CSIG - CEPS) * q.bm);
    const double invMWmix = 1.0 / q.MWmix;
    double T = (Tguess < 1.0) ? 300.0 : Tguess, cpi = 0.0, hi = 0.0;
    for (int k = 0; k < 12; k++) {
      attraction(c, &q, Y, T);
      ideal_cph(c, m, T, &cpi, &hi);
      const double app = 0.5 * (q.S2 * q.S2 - q.dam) / T;
      const double f = hi - m->R * T - (q.am - T * q.dam) * q.I * invMWmix - e;
      const double cv = cpi - m->R + T * app * q.I * invMWmix;
      const double dT = f / cv;
      T -= dT;
      if (k >= 2 && fabs(dT) <= tol * T) break;
    }
    attraction(c, &q, Y, T);
    ideal_cph(c, m, T, &cpi, &hi);
    const double vs = q.v + CSIG * q.bm, ve = q.v + CEPS * q.bm;
    const double p = c->Ru * T / (q.v - q.bm) - q.am / (vs * ve);
    dimless(c, &q, p, T);
    q.Z = p * q.invRuT * q.v;
    rootslope(&q);
    double hdep, cpdep;
    departures(c, &q, T, &hdep, &cpdep);
    s->rho = rho;
    s->rhoinv = 1.0 / rho;
    s->e = e;
    s->p = p;
    s->T = T;
    s->Z = q.Z;
    s->cp = cpi + m->R * cpdep;
    const double invZt = 1.0 / (q.Z - q.Cstar);
    gamma_c(c, &q, p, T, rho, s->cp, invZt, &s->gamma, &s->c);
    return;
  }
  double T = (Tguess < 1.0) ? 300.0 : Tguess, cp = 0.0, h = 0.0;
  for (int k = 0; k < 12; k++) {
    ideal_cph(c, m, T, &cp, &h);
    const double dT = (e - (h - m->R * T)) / (cp - m->R);
    T += dT;
    if (k >= 2 && fabs(dT) <= tol * T) break;
  }
  s->rho = rho;
  s->rhoinv = 1.0 / rho;
  s->e = e;
  s->T = T;
  s->p = rho * m->R * T;
  s->cp = cp;
  s->gamma = cp / (cp - m->R);
  s->c = sqrt(s->gamma * m->R * T);
  s->Z = 1.0;
}

/* cpg eos.mako:17-27; tpg eos.mako:22-33; real gas eos/realgas/kernels/eos.mako:137-155 */
static void frompT(const pgo_case *c, const mixmom *m, const double *Y, double p, double T, eosstate *s) {
  if (c->eos == 2) {
    cubicpt q;
    cubicmix(c, m, &q);
    attraction(c, &q, Y, T);
    double cpi, hi;
    ideal_cph(c, m, T, &cpi, &hi);
    zsolve(c, &q, p, T);
    double hdep, cpdep;
    departures(c, &q, T, &hdep, &cpdep);
    s->p = p;
    s->T = T;
    s->Z = q.Z;
    s->cp = cpi + m->R * cpdep;
    const double Zt = q.Z - q.Cstar, invZt = 1.0 / Zt, porho = Zt * m->R * T;
    s->rho = p * q.invRuT * q.MWmix * invZt;
    s->rhoinv = 1.0 / s->rho;
    s->e = hi + m->R * T * hdep - porho;
    gamma_c(c, &q, p, T, s->rho, s->cp, invZt, &s->gamma, &s->c);
    return;
  }
  s->p = p;
  s->T = T;
  s->rho = p / (m->R * T);
  s->rhoinv = 1.0 / s->rho;
  if (c->eos == 0) {
    const double cv = m->cp - m->R;
    s->e = cv * T + m->h298;
    s->cp = m->cp;
    s->gamma = m->cp / cv;
  } else {
    double cp, h;
    ideal_cph(c, m, T, &cp, &h);
    s->e = h - m->R * T;
    s->cp = cp;
    s->gamma = cp / (cp - m->R);
  }
  s->c = sqrt(s->gamma * m->R * T);
  s->Z = 1.0;
}

static void store(pgo_arr st, int i, int j, int k, const eosstate *s) {
  *at(st, i, j, k, 0) = s->p;
  *at(st, i, j, k, 1) = s->T;
  *at(st, i, j, k, 2) = s->rhoinv;
  *at(st, i, j, k, 3) = s->rho;
  *at(st, i, j, k, 4) = s->e;
  *at(st, i, j, k, 5) = s->cp;
  *at(st, i, j, k, 6) = s->gamma;
  *at(st, i, j, k, 7) = s->c;
}

/* the real gas's ninth state field */
static void storeZ(const pgo_case *c, pgo_arr st, int i, int j, int k, const eosstate *s) {
  if (c->eos == 2) *at(st, i, j, k, 8) = s->Z;
}

static void load(pgo_arr st, int i, int j, int k, eosstate *s) {
  s->p = *at(st, i, j, k, 0);
  s->T = *at(st, i, j, k, 1);
  s->rhoinv = *at(st, i, j, k, 2);
  s->rho = *at(st, i, j, k, 3);
  s->e = *at(st, i, j, k, 4);
  s->cp = *at(st, i, j, k, 5);
  s->gamma = *at(st, i, j, k, 6);
  s->c = *at(st, i, j, k, 7);
}

/* threads only where there is work for them: a region over a few cells costs more in the fork
 * and its barrier than it saves, and under a loaded host the barrier's spin is slower still */
#define PGO_PAR_CELLS 16384
#define FOR_BOX(r)                                                     \
  _Pragma("omp parallel for collapse(2) schedule(static) if ((int64_t)(r).extent[0] * (r).extent[1] * (r).extent[2] > PGO_PAR_CELLS)") \
  for (int k = (r).start[2]; k < (r).start[2] + (r).extent[2]; k++)    \
    for (int j = (r).start[1]; j < (r).start[1] + (r).extent[1]; j++)  \
      for (int i = (r).start[0]; i < (r).start[0] + (r).extent[0]; i++)

/* euler/kernels/state/fromQ.mako:5-19 with efromQ, solution.mako:10-12 */
void pgo_fromQ(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_box r) {
  FOR_BOX(r) {
    const double rho = *at(Q, i, j, k, 0), rhoinv = 1.0 / rho;
    const double q1 = *at(Q, i, j, k, 1), q2 = *at(Q, i, j, k, 2), q3 = *at(Q, i, j, k, 3);
    const double e = (*at(Q, i, j, k, 4) - 0.5 * ((q1) * (q1) + (q2) * (q2) + (q3) * (q3)) * rhoinv) * rhoinv;
    double Y[MAXNS];
    for (int n = 0; n < c->ns - 1; n++) Y[n] = *at(Q, i, j, k, 5 + n) * rhoinv;
    mixmom m;
    mix(c, Y, &m);
    eosstate s;
    fromrhoE(c, &m, Y, rho, e, *at(state, i, j, k, 1), &s);
    store(state, i, j, k, &s);
    storeZ(c, state, i, j, k, &s);
  }
}

/* euler/kernels/state/fromprims.mako:4-13 with fromprims/fromthermo, solution.mako:2-23 */
void pgo_fromprims(const pgo_case *c, pgo_arr prims, pgo_arr Q, pgo_arr state, pgo_box r) {
  FOR_BOX(r) {
    double Y[MAXNS], u[3];
    for (int n = 0; n < c->ns - 1; n++) Y[n] = *at(prims, i, j, k, 5 + n);
    for (int d = 0; d < 3; d++) u[d] = *at(prims, i, j, k, 1 + d);
    mixmom m;
    mix(c, Y, &m);
    eosstate s;
    frompT(c, &m, Y, *at(prims, i, j, k, 0), *at(prims, i, j, k, 4), &s);
    *at(Q, i, j, k, 0) = s.rho;
    *at(Q, i, j, k, 1) = s.rho * u[0];
    *at(Q, i, j, k, 2) = s.rho * u[1];
    *at(Q, i, j, k, 3) = s.rho * u[2];
    *at(Q, i, j, k, 4) = s.rho * (s.e + 0.5 * ((u[0]) * (u[0]) + (u[1]) * (u[1]) + (u[2]) * (u[2])));
    for (int n = 0; n < c->ns - 1; n++) *at(Q, i, j, k, 5 + n) = s.rho * Y[n];
    store(state, i, j, k, &s);
    storeZ(c, state, i, j, k, &s);
  }
}

/* ------------------------------------------------------------------------
 * the advective face flux: euler/kernels/flux/single.mako:4-9
 * ---------------------------------------------------------------------- */
typedef struct {
  eosstate s;
  double u, v, w;
  double Y[MAXNS];
} side;

/* recon/{koren,minmod,mc,superbee,vanleer}.mako: b phi(a/b) for b > 0 */
static double limit(int kind, double a, double b) {
  switch (kind) {
    case 1: return fmax(0.0, fmin(fmin(2.0 * a, (2.0 * b + a) * (1.0 / 3.0)), 2.0 * b));
    case 2: return fmax(0.0, fmin(b, a));
    case 3: return fmax(0.0, fmin(fmin(2.0 * a, (b + a) * 0.5), 2.0 * b));
    case 4: return fmax(0.0, fmax(fmin(2.0 * a, b), fmin(a, 2.0 * b)));
    default: return (a + fabs(a)) * b / (b + fabs(a));
  }
}

/* recon/muscl.mako:5-13 */
static double phi(int kind, double dm, double dp) {
  double p = 0.0;
  if (dp != 0.0) {
    const double s = copysign(1.0, dp), a = s * dm, b = fabs(dp);
    p = s * limit(kind, a, b);
  }
  return p;
}

double pgo_limit(int kind, double a, double b) { return limit(kind, a, b); }

static double lft(int kind, double vLL, double vL, double vR) { return vL + 0.5 * phi(kind, vL - vLL, vR - vL); }
static double rgt(int kind, double vL, double vR, double vRR) { return vR - 0.5 * phi(kind, vRR - vR, vR - vL); }

/* the face's two sides: recon/constant.mako:3-18 or recon/muscl.mako:15-51 */
static void sides(const pgo_case *c, pgo_arr Q, pgo_arr st, const int *cl, const int *d, side *L, side *R) {
  /* cl: the cell right of the face (R); d: the unit step along the axis */
  int idx[4][3];
  for (int o = 0; o < 4; o++)
    for (int a = 0; a < 3; a++) idx[o][a] = cl[a] + (o - 2) * d[a];
#define QQ(o, comp) (*at(Q, idx[o][0], idx[o][1], idx[o][2], comp))
#define SS(o, f) (*at(st, idx[o][0], idx[o][1], idx[o][2], f))
  const int ns = c->ns;
  if (c->recon == 0) {
    load(st, idx[1][0], idx[1][1], idx[1][2], &L->s);
    load(st, idx[2][0], idx[2][1], idx[2][2], &R->s);
    L->u = QQ(1, 1) / QQ(1, 0);
    L->v = QQ(1, 2) / QQ(1, 0);
    L->w = QQ(1, 3) / QQ(1, 0);
    R->u = QQ(2, 1) / QQ(2, 0);
    R->v = QQ(2, 2) / QQ(2, 0);
    R->w = QQ(2, 3) / QQ(2, 0);
    for (int n = 0; n < ns - 1; n++) {
      L->Y[n] = QQ(1, 5 + n) / QQ(1, 0);
      R->Y[n] = QQ(2, 5 + n) / QQ(2, 0);
    }
    return;
  }
  const int lim = c->limiter;
  const double ri[4] = {SS(0, 2), SS(1, 2), SS(2, 2), SS(3, 2)};
  const double pL = lft(lim, SS(0, 0), SS(1, 0), SS(2, 0)), pR = rgt(lim, SS(1, 0), SS(2, 0), SS(3, 0));
  const double TL = lft(lim, SS(0, 1), SS(1, 1), SS(2, 1)), TR = rgt(lim, SS(1, 1), SS(2, 1), SS(3, 1));
  for (int n = 0; n < ns - 1; n++) {
    L->Y[n] = lft(lim, QQ(0, 5 + n) * ri[0], QQ(1, 5 + n) * ri[1], QQ(2, 5 + n) * ri[2]);
    R->Y[n] = rgt(lim, QQ(1, 5 + n) * ri[1], QQ(2, 5 + n) * ri[2], QQ(3, 5 + n) * ri[3]);
  }
  mixmom m;
  mix(c, L->Y, &m);
  frompT(c, &m, L->Y, pL, TL, &L->s);
  L->u = lft(lim, QQ(0, 1) * ri[0], QQ(1, 1) * ri[1], QQ(2, 1) * ri[2]);
  L->v = lft(lim, QQ(0, 2) * ri[0], QQ(1, 2) * ri[1], QQ(2, 2) * ri[2]);
  L->w = lft(lim, QQ(0, 3) * ri[0], QQ(1, 3) * ri[1], QQ(2, 3) * ri[2]);
  mix(c, R->Y, &m);
  frompT(c, &m, R->Y, pR, TR, &R->s);
  R->u = rgt(lim, QQ(1, 1) * ri[1], QQ(2, 1) * ri[2], QQ(3, 1) * ri[3]);
  R->v = rgt(lim, QQ(1, 2) * ri[1], QQ(2, 2) * ri[2], QQ(3, 2) * ri[3]);
  R->w = rgt(lim, QQ(1, 3) * ri[1], QQ(2, 3) * ri[2], QQ(3, 3) * ri[3]);
#undef QQ
#undef SS
}

/* the two sides of the face whose right cell is (i, j, k) along axis, each written as
 * p, T, rho, e, c, u, v, w, Y_0 .. Y_{ns-2} (left first): the reconstruction pgo_flux feeds its
 * flux function */
void pgo_sides(const pgo_case *c, pgo_arr Q, pgo_arr st, int axis, int i, int j, int k, double *out) {
  const int cl[3] = {i, j, k};
  int d[3] = {0, 0, 0};
  d[axis] = 1;
  side L, R;
  sides(c, Q, st, cl, d, &L, &R);
  const side *X[2] = {&L, &R};
  const int w = 8 + c->ns - 1;
  for (int s = 0; s < 2; s++) {
    double *o = out + s * w;
    o[0] = X[s]->s.p, o[1] = X[s]->s.T, o[2] = X[s]->s.rho, o[3] = X[s]->s.e, o[4] = X[s]->s.c;
    o[5] = X[s]->u, o[6] = X[s]->v, o[7] = X[s]->w;
    for (int n = 0; n < c->ns - 1; n++) o[8 + n] = X[s]->Y[n];
  }
}

/* function/kepaec.mako:2-20 (Shima et al. 2021) */
static void kepaec(int ns, const side *L, const side *R, const double *S, double *F) {
  const double uf = 0.5 * (R->u + L->u), vf = 0.5 * (R->v + L->v), wf = 0.5 * (R->w + L->w);
  const double U = ((S[0]) * (uf) + (S[1]) * (vf) + (S[2]) * (wf));
  const double pf = 0.5 * (R->s.p + L->s.p);
  const double rho = 0.5 * (R->s.rho + L->s.rho);
  const double C = rho * U;
  F[0] = C;
  F[1] = C * uf + pf * S[0];
  F[2] = C * vf + pf * S[1];
  F[3] = C * wf + pf * S[2];
  const double Kj = C * 0.5 * ((R->u) * (L->u) + (R->v) * (L->v) + (R->w) * (L->w));
  const double Pj = 0.5 * (L->s.p * ((R->u) * (S[0]) + (R->v) * (S[1]) + (R->w) * (S[2])) +
                           R->s.p * ((L->u) * (S[0]) + (L->v) * (S[1]) + (L->w) * (S[2])));
  const double thL = L->s.cp / L->s.gamma * L->s.T, thR = R->s.cp / R->s.gamma * R->s.T;
  const double ef = 2.0 * thL * thR / (thL + thR) + 0.5 * ((L->s.e - thL) + (R->s.e - thR));
  F[4] = ef * C + Kj + Pj;
  for (int n = 0; n < ns - 1; n++) F[5 + n] = 0.5 * (R->Y[n] + L->Y[n]) * C;
}

/* function/hllc.mako:2-43 */
static void hllc(int ns, const side *L, const side *R, const double *S, double *F) {
  const double area = sqrt(((S[0]) * (S[0]) + (S[1]) * (S[1]) + (S[2]) * (S[2])));
  const double nx = S[0] / area, ny = S[1] / area, nz = S[2] / area;
  const double unL = ((L->u) * (S[0]) + (L->v) * (S[1]) + (L->w) * (S[2])) / area;
  const double unR = ((R->u) * (S[0]) + (R->v) * (S[1]) + (R->w) * (S[2])) / area;
  const double SL = unL - L->s.c, SR = unR + R->s.c;
  const double Sstar = (R->s.p - L->s.p + L->s.rho * unL * (SL - unL) - R->s.rho * unR * (SR - unR)) /
                       (L->s.rho * (SL - unL) - R->s.rho * (SR - unR));
  int left, star;
  if (SL >= 0.0) {
    left = 1, star = 0;
  } else if (Sstar >= 0.0) {
    left = 1, star = 1;
  } else if (SR >= 0.0) {
    left = 0, star = 1;
  } else {
    left = 0, star = 0;
  }
  const side *X = left ? L : R;
  const double U = left ? unL : unR, SX = left ? SL : SR;
  const double et = X->s.e + 0.5 * ((X->u) * (X->u) + (X->v) * (X->v) + (X->w) * (X->w));
  const double mdot = U * X->s.rho * area;
  F[0] = mdot;
  F[1] = mdot * X->u + X->s.p * S[0];
  F[2] = mdot * X->v + X->s.p * S[1];
  F[3] = mdot * X->w + X->s.p * S[2];
  F[4] = mdot * et + U * X->s.p * area;
  for (int n = 0; n < ns - 1; n++) F[5 + n] = mdot * X->Y[n];
  if (star) {
    const double rhostar = X->s.rho * ((SX - U) / (SX - Sstar));
    const double drho = rhostar - X->s.rho, dU = rhostar * (Sstar - U);
    const double jump = SX * area;
    F[0] += jump * drho;
    F[1] += jump * (drho * X->u + dU * nx);
    F[2] += jump * (drho * X->v + dU * ny);
    F[3] += jump * (drho * X->w + dU * nz);
    F[4] += jump * (drho * et + dU * (Sstar + X->s.p / (X->s.rho * (SX - U))));
    for (int n = 0; n < ns - 1; n++) F[5 + n] += jump * drho * X->Y[n];
  }
}

/* function/rusanov.mako:2-19 */
static void rusanov(int ns, const side *L, const side *R, const double *S, double *F) {
  const double area = sqrt(((S[0]) * (S[0]) + (S[1]) * (S[1]) + (S[2]) * (S[2])));
  const double unL = ((L->u) * (S[0]) + (L->v) * (S[1]) + (L->w) * (S[2])) / area;
  const double unR = ((R->u) * (S[0]) + (R->v) * (S[1]) + (R->w) * (S[2])) / area;
  const double lam = fmax(fabs(unL) + L->s.c, fabs(unR) + R->s.c) * area;
  const double UL = unL * area, UR = unR * area;
  const double rL = L->s.rho, rR = R->s.rho;
  F[0] = 0.5 * (UR * rR + UL * rL - lam * (rR - rL));
  F[1] = 0.5 * (UR * rR * R->u + R->s.p * S[0] + UL * rL * L->u + L->s.p * S[0] - lam * (rR * R->u - rL * L->u));
  F[2] = 0.5 * (UR * rR * R->v + R->s.p * S[1] + UL * rL * L->v + L->s.p * S[1] - lam * (rR * R->v - rL * L->v));
  F[3] = 0.5 * (UR * rR * R->w + R->s.p * S[2] + UL * rL * L->w + L->s.p * S[2] - lam * (rR * R->w - rL * L->w));
  const double EL = rL * (L->s.e + 0.5 * ((L->u) * (L->u) + (L->v) * (L->v) + (L->w) * (L->w)));
  const double ER = rR * (R->s.e + 0.5 * ((R->u) * (R->u) + (R->v) * (R->v) + (R->w) * (R->w)));
  F[4] = 0.5 * (UR * (ER + R->s.p) + UL * (EL + L->s.p) - lam * (ER - EL));
  for (int n = 0; n < ns - 1; n++) {
    const double rhoYL = rL * L->Y[n], rhoYR = rR * R->Y[n];
    F[5 + n] = 0.5 * (rhoYR * UR + rhoYL * UL - lam * (rhoYR - rhoYL));
  }
}

void pgo_flux(const pgo_case *c, int axis, pgo_arr Q, pgo_arr state, pgo_arr S, pgo_arr F, pgo_box r) {
  int d[3] = {0, 0, 0};
  d[axis] = 1;
  FOR_BOX(r) {
    side L, R;
    const int cl[3] = {i, j, k};
    sides(c, Q, state, cl, d, &L, &R);
    const double Sv[3] = {*at(S, i, j, k, 0), *at(S, i, j, k, 1), *at(S, i, j, k, 2)};
    double Fv[5 + MAXNS];
    if (c->primary == 0)
      kepaec(c->ns, &L, &R, Sv, Fv);
    else if (c->primary == 1)
      hllc(c->ns, &L, &R, Sv, Fv);
    else
      rusanov(c->ns, &L, &R, Sv, Fv);
    for (int l = 0; l < c->ne; l++) *at(F, i, j, k, l) = Fv[l];
  }
}

/* advec/kernels/applyflux.mako:2-11, summed low in, then high out */
void pgo_applyflux(int ne, pgo_arr dQ, pgo_arr iF, pgo_arr jF, pgo_arr kF, pgo_arr detJinv, int append, pgo_box r) {
  FOR_BOX(r) {
    const double J = *at(detJinv, i, j, k, 0);
    for (int l = 0; l < ne; l++) {
      const double v = (*at(iF, i, j, k, l) + *at(jF, i, j, k, l) + *at(kF, i, j, k, l) - *at(iF, i + 1, j, k, l) -
                        *at(jF, i, j + 1, k, l) - *at(kF, i, j, k + 1, l)) * J;
      if (append)
        *at(dQ, i, j, k, l) += v;
      else
        *at(dQ, i, j, k, l) = v;
    }
  }
}

/* ------------------------------------------------------------------------
 * finite-rate production: chemistry/kernels/finiterate.mako, table-driven
 * ---------------------------------------------------------------------- */
static double troeblend(const pgo_case *c, int k, double logT, double logpr) {
  const double log10e = 0.4342944819032518, ln10 = 2.302585092994046;
  const double *P = c->lfcpoly + 13 * c->troeidx[k];
  const double tf = (logT - c->troeuc) * c->troeuhinv;
  double lfc = P[12];
  for (int m = 11; m >= 0; m--) lfc = lfc * tf + P[m];
  const double C = -0.4 - 0.67 * lfc, N = 0.75 - 1.27 * lfc;
  const double A = logpr * log10e + C, B = N - 0.14 * A, B2 = B * B;
  return lfc * B2 / (B2 + A * A) * ln10;
}

/* ln(1 + x) through elog: u = 1 + x rounded, then the first-order correction for that rounding,
 * ((u - 1) - x) / u. One specified method on both sides, in place of the libraries' log1p, whose
 * last bits differ between libm and the device */
double pgo_dlog1p(double x) {
  const double u = 1.0 + x;
  return pgo_elog(u) - ((u - 1.0) - x) / u;
}

/* omega, and the destruction rates D unless D is null: one fold for both (finiterate.mako:185-236) */
void pgo_sources(const pgo_case *c, double rho, const double *Y, double T, double *omega, double *D) {
  const int ns = c->ns;
  /* ratestate, finiterate.mako:136-182 */
  const double logT = pgo_tlog(T), Tinv = 1.0 / T;
  const double logprefRuT = c->logprefRu - logT;
  const double prefRuT = c->prefRu * Tinv, prefRuTinv = c->Rupref * T;
  double ctot = 0.0, gRT[MAXNS], logc[MAXNS], cG[MAXNS], conc[MAXNS];
  for (int n = 0; n < ns; n++) {
    const double cn = rho * Y[n] * c->MWinv[n];
    conc[n] = cn;
    ctot += cn;
    double g = 0.0;
    for (int m = c->ng_ - 1; m >= 0; m--) g = FMA(g, logT, c->gpoly[n * c->ng_ + m]);
    gRT[n] = FMA(c->h298[n], Tinv, g);
    logc[n] = fmax(c->logcfloor, pgo_tlog(cn));
    cG[n] = fmax(cn, 0.0) * pgo_texp(gRT[n]);
  }
  double ctb[MAXTB], logtb[MAXTB];
  for (int j = 0; j < c->ntb; j++) {
    double v = c->tbdefeff[j] * ctot;
    for (int e = 0; e < c->ntbeff[j]; e++) {
      const int sp = c->tbeffsp[j * ns + e];
      v += c->tbeffdev[j * ns + e] * (rho * Y[sp] * c->MWinv[sp]);
    }
    ctb[j] = v;
    logtb[j] = pgo_tlog(fmax(DBL_MIN, v));
  }
  (void)conc;
  /* fold, finiterate.mako:185-226 */
  const double expclamp = c->expclamp;
  for (int n = 0; n < ns; n++) omega[n] = 0.0;
  if (D)
    for (int n = 0; n < ns; n++) D[n] = 0.0;
  for (int k = 0; k < c->nr; k++) {
    const int kind = c->kind[k], thirdbody = kind >= 1, falloff = kind >= 2, troe = kind == 3;
    const int tb = c->tb[k];
    double rp, fwd, rev;
    double lk = c->logA[k];
    if (c->b[k] != 0.0) lk += c->b[k] * logT;
    if (c->EaR[k] != 0.0) lk -= c->EaR[k] * Tinv;
    if (c->product[k]) {
      /* kform, finiterate.mako:69-125 */
      double logpr = 0.0, pfac = 1.0;
      if (falloff) {
        logpr = logtb[tb] + c->logprA[k];
        if (c->prb[k] != 0.0) logpr += c->prb[k] * logT;
        if (c->prEaR[k] != 0.0) logpr -= c->prEaR[k] * Tinv;
        const double pr = pgo_texp(fmin(logpr, expclamp));
        pfac = pr / (1.0 + pr);
      }
      if (troe) lk += troeblend(c, k, logT, logpr);
      double xk = lk;
      for (int j = 0; j < c->nreact[k]; j++) xk -= c->reactnu[4 * k + j] * gRT[c->reactsp[4 * k + j]];
      double K = pgo_texp(fmin(xk, expclamp));
      if (thirdbody && !falloff) K *= ctb[tb];
      if (falloff) K *= pfac;
      fwd = K;
      for (int j = 0; j < c->nreact[k]; j++)
        for (int t = 0; t < (int)c->reactnu[4 * k + j]; t++) fwd *= cG[c->reactsp[4 * k + j]];
      rev = 0.0;
      if (c->reversible[k]) {
        rev = K;
        const double nus = c->nusum[k];
        for (int t = 0; t < (int)fabs(nus); t++) rev *= nus > 0 ? prefRuTinv : prefRuT;
        for (int j = 0; j < c->nprod[k]; j++)
          for (int t = 0; t < (int)c->prodnu[4 * k + j]; t++) rev *= cG[c->prodsp[4 * k + j]];
      }
      rp = fwd - rev;
    } else {
      /* progress, finiterate.mako:7-48 */
      if (thirdbody && !falloff) lk += logtb[tb];
      double logpr = 0.0;
      if (falloff) {
        logpr = logtb[tb] + c->logprA[k];
        if (c->prb[k] != 0.0) logpr += c->prb[k] * logT;
        if (c->prEaR[k] != 0.0) logpr -= c->prEaR[k] * Tinv;
        const double prexp = pgo_texp(-fabs(logpr));
        lk += fmin(logpr, 0.0) - DLOG1P(prexp);
      }
      if (troe) lk += troeblend(c, k, logT, logpr);
      double lrp = lk;
      for (int j = 0; j < c->nfwd[k]; j++) lrp += c->fwdorder[4 * k + j] * logc[c->fwdsp[4 * k + j]];
      if (c->reversible[k]) {
        double diff = c->nusum[k] != 0.0 ? c->nusum[k] * logprefRuT : 0.0;
        for (int j = 0; j < c->nnet[k]; j++) {
          const int sp = c->netsp[4 * k + j];
          diff -= c->netnu[4 * k + j] * (gRT[sp] + logc[sp]);
        }
        fwd = pgo_texp(fmin(lrp, expclamp));
        rev = pgo_texp(fmin(lrp - diff, expclamp));
        rp = fwd - rev;
      } else {
        fwd = pgo_texp(fmin(lrp, expclamp));
        rev = 0.0;
        rp = fwd;
      }
    }
    for (int j = 0; j < c->nnet[k]; j++) omega[c->netsp[4 * k + j]] += c->netnuMW[4 * k + j] * rp;
    if (D)
      for (int j = 0; j < c->nnet[k]; j++) {
        const double nuMW = c->netnuMW[4 * k + j];
        D[c->netsp[4 * k + j]] += (nuMW < 0.0 ? -nuMW : 0.0) * fwd + (nuMW > 0.0 ? nuMW : 0.0) * rev;
      }
  }
}

void pgo_production(const pgo_case *c, double rho, const double *Y, double T, double *omega) {
  pgo_sources(c, rho, Y, T, omega, NULL);
}

/* chemint/kernels/source.mako:3-10, cell.mako:5-20, massfrac solution.mako:25-35 */
void pgo_source(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr dQ, pgo_box r) {
  const int ns = c->ns;
  FOR_BOX(r) {
    const double rho = *at(Q, i, j, k, 0), rhoinv = 1.0 / rho;
    double Ylast = 1.0, Ysum = 0.0, Y[MAXNS], omega[MAXNS];
    for (int n = 0; n < ns - 1; n++) {
      const double Yn = fmax(fmin(*at(Q, i, j, k, 5 + n) * rhoinv, 1.0), 0.0);
      Ylast -= Yn;
      Ysum += Yn;
    }
    const double Yscale = 1.0 / fmax(Ysum, 1.0);
    Ylast = fmax(Ylast, 0.0);
    for (int n = 0; n < ns - 1; n++) Y[n] = fmax(fmin(*at(Q, i, j, k, 5 + n) * rhoinv, 1.0), 0.0) * Yscale;
    Y[ns - 1] = Ylast;
    pgo_production(c, rho, Y, *at(state, i, j, k, 1), omega);
    for (int l = 0; l < 5; l++) *at(dQ, i, j, k, l) = 0.0;
    for (int n = 0; n < ns - 1; n++) *at(dQ, i, j, k, 5 + n) = omega[n];
  }
}

/* integrators/kernels/combine.mako:1-13: out = [wout out] + w0 X0 + ... + w_last dt X_last */
void pgo_combine(int ne, pgo_arr out, int own, double wout, int nterms, const pgo_arr *X, const double *w,
                 double dt, int step, pgo_box r) {
  FOR_BOX(r) {
    for (int l = 0; l < ne; l++) {
      double acc = 0.0;
      int first = 1;
      if (own) {
        acc = wout * *at(out, i, j, k, l);
        first = 0;
      }
      for (int m = 0; m < nterms; m++) {
        const double term = (step && m == nterms - 1) ? w[m] * dt * *at(X[m], i, j, k, l) : w[m] * *at(X[m], i, j, k, l);
        acc = first ? term : acc + term;
        first = 0;
      }
      *at(out, i, j, k, l) = acc;
    }
  }
}

/* multiblock/collections/exchange/kernels/place.mako:11-37 on a block periodic with itself:
 * each halo slab takes the interior planes next to the opposite interface; the other
 * axes run over the interior only (edges and corners are never filled) */
void pgo_fill_periodic(pgo_arr A, int ncomp, const int *ncell, int ng, int depth, int axis) {
  const int n = ncell[axis];
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = ng;
    hi[a] = ng + ncell[a];
  }
  lo[axis] = 0;
  hi[axis] = depth;
#pragma omp parallel for collapse(2) schedule(static) if ((int64_t)ncomp * (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]) > PGO_PAR_CELLS)
  for (int l = 0; l < ncomp; l++)
    for (int k = lo[2]; k < hi[2]; k++)
      for (int j = lo[1]; j < hi[1]; j++)
        for (int i = lo[0]; i < hi[0]; i++) {
          int h[3] = {i, j, k};
          /* layer h[axis] = 0..depth-1 counted outward */
          const int layer = h[axis];
          int dlo[3] = {i, j, k}, slo[3] = {i, j, k}, dhi[3] = {i, j, k}, shi[3] = {i, j, k};
          dlo[axis] = ng - 1 - layer;
          slo[axis] = ng - 1 - layer + n;
          dhi[axis] = ng + n + layer;
          shi[axis] = ng + layer;
          *at(A, dlo[0], dlo[1], dlo[2], l) = *at(A, slo[0], slo[1], slo[2], l);
          *at(A, dhi[0], dhi[1], dhi[2], l) = *at(A, shi[0], shi[1], shi[2], l);
        }
}

/* ------------------------------------------------------------------------
 * a whole stage and step: multiblock/solver.py:246-284,355-364,
 * simulator/advec/simulator.py:57-87, integrators/rungekutta.py:9-95
 * ---------------------------------------------------------------------- */
/* ------------------------------------------------------------------------
 * the viscous terms
 * ---------------------------------------------------------------------- */

/* mixrule/kernels/wilke.mako:3-17: phi_nm = (1 + s_n / s_m)^2 wilkepair(n, m), with
 * s_n = sqrt(mu_n) MW_n^(-1/4) */
static double viscmix(const pgo_case *c, const double *X, const double *sqrtmu) {
  const int ns = c->ns;
  double sinv[MAXNS];
  for (int n = 0; n < ns; n++) sinv[n] = 1.0 / (sqrtmu[n] * c->MWqinv[n]);
  double mu = 0.0;
  for (int n = 0; n < ns; n++) {
    const double sn = sqrtmu[n] * c->MWqinv[n];
    double phi = 0.0;
    for (int m = 0; m < ns; m++) {
      const double r = 1.0 + sn * sinv[m];
      phi += r * r * c->wilkepair[n * ns + m] * X[m];
    }
    mu += sqrtmu[n] * sqrtmu[n] * X[n] / phi;
  }
  return mu;
}

/* transport/kernels/mixing.mako:3-10: the mean of the series and parallel sums */
static double kappamix(const pgo_case *c, const double *X, const double *kappasp) {
  double series = 0.0, parallel = 0.0;
  for (int n = 0; n < c->ns; n++) {
    series += X[n] * kappasp[n];
    parallel += X[n] / kappasp[n];
  }
  return 0.5 * (series + 1.0 / parallel);
}

/* transport/kernels/chung.mako:6-67: Chung et al.'s dense-gas factors on the dilute values,
 * one-fluid mixing, no dipole or association terms; rho in kg/m3, Chung's units inside */
static void chung(const pgo_case *c, const double *X, double T, double lnT, double rho, double *mu, double *kappa) {
  const int ns = c->ns;
  double sigma3 = 0.0, epssigma3 = 0.0, masssum = 0.0, omegasigma3 = 0.0, cvR = 0.0;
  for (int i = 0; i < ns; i++) {
    double cpR = 0.0;
    for (int m = c->nt - 1; m >= 0; m--) cpR = cpR * lnT + c->cppoly[i * c->nt + m];
    cvR += X[i] * (cpR - 1.0);
    for (int j = 0; j < ns; j++) {
      const double xx = X[i] * X[j];
      const double sigma2ij = c->chungsigma[i] * c->chungsigma[j];
      const double sigma3ij = sigma2ij * sqrt(sigma2ij);
      const double epsij = sqrt(c->chungepsk[i] * c->chungepsk[j]);
      const double Mij = 2.0 * c->MW[i] * c->MW[j] / (c->MW[i] + c->MW[j]);
      sigma3 += xx * sigma3ij;
      epssigma3 += xx * epsij * sigma3ij;
      masssum += xx * epsij * sigma2ij * sqrt(Mij);
      omegasigma3 += xx * 0.5 * (c->acentric[i] + c->acentric[j]) * sigma3ij;
    }
  }
  const double epsk = epssigma3 / sigma3;
  const double cbrtsigma3 = DCBRT(sigma3);
  const double sigma2 = cbrtsigma3 * cbrtsigma3;
  const double sqrtMm = masssum / (epsk * sigma2);
  const double Mm = sqrtMm * sqrtMm;
  const double omega = omegasigma3 / sigma3;
  const double Tc = 1.2593 * epsk;
  const double Vc = sigma3 / (0.809 * 0.809 * 0.809);
  const double Fc = 1.0 - 0.2756 * omega;
  const double cbrtVc = DCBRT(Vc);
  const double Vc23 = cbrtVc * cbrtVc;
  const double Mprime = Mm * 1.0e-3;
  double E[10], B[7];
  for (int i = 0; i < 10; i++) E[i] = c->chungvisc[4 * i] + c->chungvisc[4 * i + 1] * omega;
  for (int i = 0; i < 7; i++) B[i] = c->chungcond[4 * i] + c->chungcond[4 * i + 1] * omega;
  const double Tr = T / Tc;
  const double Tstar = T / epsk, invTstar = epsk / T;
  /* pow and exp by the table methods (the device evaluates the same) */
  const double Omegav = 1.16145 * pgo_texp(-0.14874 * ELOG(Tstar)) + 0.52487 * pgo_texp(-0.77320 * Tstar) +
                        2.16178 * pgo_texp(-2.43787 * Tstar);
  const double Yy = rho / Mm * 1.0e-3 * Vc / 6.0;
  const double omYy = 1.0 - Yy;
  const double G1 = (1.0 - 0.5 * Yy) / (omYy * omYy * omYy);
  const double G2 = (E[0] * (1.0 - pgo_texp(-E[3] * Yy)) / Yy + E[1] * G1 * pgo_texp(E[4] * Yy) + E[2] * G1) /
                    (E[0] * E[3] + E[1] + E[2]);
  const double etastarstar = E[6] * Yy * Yy * G2 * pgo_texp(E[7] + E[8] * invTstar + E[9] * invTstar * invTstar);
  const double muratio = (1.0 / G2 + E[5] * Yy) + etastarstar * Omegav / (sqrt(Tstar) * Fc);
  *mu = *mu * muratio;
  const double eta0 = 4.0785e-6 * sqrt(Mm * T) / (Vc23 * Omegav) * Fc;
  const double alpha = cvR - 1.5;
  const double beta = 0.7862 - 0.7109 * omega + 1.3168 * omega * omega;
  const double Zc = 2.0 + 10.5 * Tr * Tr;
  const double Psi = 1.0 + alpha * ((0.215 + 0.28288 * alpha - 1.061 * beta + 0.26665 * Zc) / (0.6366 + beta * Zc + 1.061 * alpha * beta));
  const double lambda0 = 31.2 * eta0 * Psi / Mprime;
  const double H2 = (B[0] * (1.0 - pgo_texp(-B[3] * Yy)) / Yy + B[1] * G1 * pgo_texp(B[4] * Yy) + B[2] * G1) /
                    (B[0] * B[3] + B[1] + B[2]);
  const double kapparatio = (1.0 / H2 + B[5] * Yy) + 3.586e-3 * sqrt(Tc / Mprime) / Vc23 * B[6] * Yy * Yy * H2 * sqrt(Tr) / lambda0;
  *kappa = *kappa * kapparatio;
}

/* eos/realgas/kernels/eos.mako:192-273 hspecies: the partial molar enthalpies of the cubic,
 * -Ru T^2 d ln phi_n / dT plus the ideal part and the volume translation's -p c_n, per mass,
 * at the state's own root Z */
static void hpartial(const pgo_case *c, const double *Yall, double p, double T, double Z, double *h) {
  const int ns = c->ns;
  mixmom m;
  mix(c, Yall, &m);
  cubicpt q;
  cubicmix(c, &m, &q);
  attraction(c, &q, Yall, T);
  dimless(c, &q, p, T);
  q.Z = Z;
  rootslope(&q);
  double hdep, cpdep;
  departures(c, &q, T, &hdep, &cpdep);
  const double lnT = ELOG(T), Tinv = c->Ru * q.invRuT, sqrtT = sqrt(T);
  const double ZT = q.dZdT, Bs = q.Bstar, BT = -Bs * Tinv;
  const double invZsB = 1.0 / (Z + CSIG * Bs), invZeB = 1.0 / (Z + CEPS * Bs);
  const double Iz = -q.logratio / (CSIG - CEPS);
  const double IzT = ((ZT + CSIG * BT) * invZsB - (ZT + CEPS * BT) * invZeB) / (CSIG - CEPS);
  const double dlnZmB = (ZT - BT) / (Z - Bs);
  const double invS1 = 1.0 / q.S1, invBm = 1.0 / q.bm, invamsqrtT = invS1 * invS1 / sqrtT;
  for (int n = 0; n < ns; n++) {
    double hRT = 0.0;
    for (int mm = c->nt - 1; mm >= 0; mm--) hRT = FMA(hRT, lnT, c->hpoly[n * c->nt + mm]);
    hRT += c->h298[n] * Tinv;
    const double qT = (q.qd - q.q) * Tinv;
    const double sqrtan = c->cubicsqrtac[n] * modulus(c, n, sqrtT);
    const double r = 2.0 * sqrtan * invS1;
    const double rT = (sqrtan * q.S2 - c->cubicda[n] * q.S1) * invamsqrtT;
    const double bb = c->cubicb[n] * invBm;
    const double qbar = q.q * (r - bb), qbarT = qT * (r - bb) + q.q * rT;
    const double dlnphidT = bb * q.dZdT - dlnZmB - qbarT * Iz - qbar * IzT;
    h[n] = (hRT * T * c->Ru - c->Ru * T * T * dlnphidT - c->vshift[n] * p) * c->MWinv[n];
  }
}

/* the mixture's mu (returned) and kappa at T (lnT its log) and rho: constant species values or the
 * kinetic-theory fits (transport/kernels/{constant,dilute}.mako), Wilke's rule and the
 * conductivity mean (mixrule/kernels/wilke.mako:3-17, transport/kernels/mixing.mako:3-10), Chung's
 * dense factors (transport/kernels/chung.mako) */
static double mu_kappa(const pgo_case *c, const double *X, double T, double lnT, double rho, double *kappa) {
  const int ns = c->ns;
  double sqrtmu[MAXNS], kappasp[MAXNS];
  if (c->trans == 0) {
    for (int n = 0; n < ns; n++) sqrtmu[n] = sqrt(c->mu0[n]), kappasp[n] = c->kappa0[n];
  } else {
    /* dilute.mako:5-26: sqrt(mu) = fit(ln T) T^(1/4), kappa = fit(ln T) sqrt(T) */
    const double sqrtT = sqrt(T), sqrtsqrtT = sqrt(sqrtT);
    for (int n = 0; n < ns; n++) {
      double fit = 0.0;
      for (int m = c->nmu - 1; m >= 0; m--) fit = fit * lnT + c->mupoly[n * c->nmu + m];
      sqrtmu[n] = fit * sqrtsqrtT;
    }
    for (int n = 0; n < ns; n++) {
      double fit = 0.0;
      for (int m = c->nkappa - 1; m >= 0; m--) fit = fit * lnT + c->kappapoly[n * c->nkappa + m];
      kappasp[n] = fit * sqrtT;
    }
  }
  double mu = viscmix(c, X, sqrtmu);
  *kappa = kappamix(c, X, kappasp);
  if (c->trans == 2) chung(c, X, T, lnT, rho, &mu, kappa);
  return mu;
}

/* diffusion/kernels/binary.mako:11-62: each species' mixture-averaged coefficient over the bath,
 * D_n = rest / (p (rest sum1 + X_n sum2)) with rest = sum X_m MW_m, sum1 = sum X_m / D_nm and
 * sum2 = sum X_m MW_m / D_nm over the partners (a partial bath scales the first sum from the
 * bath's fraction to the whole); a species alone takes its self-diffusion coefficient */
static void binary_diffusion(const pgo_case *c, const double *X, double p, double T, double lnT, double *D) {
  const int ns = c->ns, np = c->npair, width = c->diffwidth, partial = c->diffpartial;
  const double invT32 = 1.0 / (T * sqrt(T)), tiny = 64.0 * DBL_EPSILON;
  for (int n = 0; n < ns; n++) {
    double sum1 = 0.0, sum2 = 0.0, rest = 0.0, bathx = 0.0;
    for (int k = 0; k < width; k++) {
      int m;
      double Xm;
      if (partial) {
        m = (int)c->bathof[n * ns + k];
        Xm = m != n ? fmax(X[m], 0.0) : 0.0;
      } else {
        m = k + (k >= n);
        Xm = fmax(X[m], 0.0);
      }
      double Dinv = 0.0;
      for (int t = np - 1; t >= 0; t--) Dinv = Dinv * lnT + c->pairfit[(n * ns + k) * np + t];
      Dinv *= invT32;
      sum1 += Xm * Dinv;
      sum2 += Xm * c->MW[m] * Dinv;
      rest += Xm * c->MW[m];
      bathx += Xm;
    }
    double Dinvself = 0.0;
    for (int t = np - 1; t >= 0; t--) Dinvself = Dinvself * lnT + c->selffit[n * np + t];
    Dinvself *= invT32;
    const double Xn = fmax(X[n], 0.0);
    int alone;
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

void pgo_viscsponge(pgo_arr xcc, pgo_arr trstate, const double *line, pgo_box r) {
  const double nx = line[0], ny = line[1], nz = line[2], start = line[3], length = line[4], m = line[5];
  FOR_BOX(r) {
    const double along = (*at(xcc, i, j, k, 0) * nx + *at(xcc, i, j, k, 1) * ny + *at(xcc, i, j, k, 2) * nz - start) /
                         length;
    *at(trstate, i, j, k, 0) *= 1.0 + fmin(fmax(0.0, along * (m - 1.0)), m - 1.0);
  }
}

/* navstokes/kernels/transport.mako:5-35 with transport/kernels/{constant,kinetic,dilute}.mako,
 * diffusion/kernels/{lewis,binary}.mako and the ideal hspecies (eos/kernels/species.mako:29-36).
 * ln T is taken by the EOS's log (pgo_elog), as everywhere in this path. */
void pgo_transport(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr trstate, pgo_arr hsp, pgo_box r) {
  const int ns = c->ns;
  FOR_BOX(r) {
    const double rho = *at(Q, i, j, k, 0), rhoinv = 1.0 / rho;
    double Y[MAXNS];
    double Ylast = 1.0;
    for (int n = 0; n < ns - 1; n++) {
      Y[n] = *at(Q, i, j, k, 5 + n) * rhoinv;
      Ylast -= *at(Q, i, j, k, 5 + n) * rhoinv;
    }
    Y[ns - 1] = fmax(Ylast, 0.0);
    /* eos/kernels/mixing.mako:34-39 molmass */
    double denom = 0.0;
    for (int n = 0; n < ns; n++) denom += Y[n] * c->MWinv[n];
    const double MWmix = 1.0 / denom;
    double X[MAXNS];
    for (int n = 0; n < ns; n++) X[n] = Y[n] * c->MWinv[n] * MWmix;
    const double T = *at(state, i, j, k, 1), lnT = ELOG(T);
    double kappa;
    const double mu = mu_kappa(c, X, T, lnT, rho, &kappa);
    *at(trstate, i, j, k, 0) = mu;
    *at(trstate, i, j, k, 1) = kappa;
    if (c->diffusion == 1) {
      double D[MAXNS];
      binary_diffusion(c, X, *at(state, i, j, k, 0), T, lnT, D);
      for (int n = 0; n < ns; n++) *at(trstate, i, j, k, 2 + n) = D[n];
    } else {
      const double kappaoverrhocp = kappa * rhoinv / *at(state, i, j, k, 5);
      for (int n = 0; n < ns; n++) *at(trstate, i, j, k, 2 + n) = kappaoverrhocp / c->lewis[n];
    }
    if (c->eos == 2) {
      double h[MAXNS];
      hpartial(c, Y, *at(state, i, j, k, 0), T, *at(state, i, j, k, 8), h);
      for (int n = 0; n < ns; n++) *at(hsp, i, j, k, n) = h[n];
      continue;
    }
    const double Tinv = 1.0 / T;
    for (int n = 0; n < ns; n++) {
      double hRT = 0.0;
      for (int m = c->nt - 1; m >= 0; m--) hRT = FMA(hRT, lnT, c->hpoly[n * c->nt + m]);
      hRT += c->h298[n] * Tinv;
      *at(hsp, i, j, k, n) = hRT * T * c->Ru * c->MWinv[n];
    }
  }
}

/* navstokes/kernels/dqdxyz.mako:16-41: central differences in index space mapped by Jinv;
 * a velocity or mass fraction as the difference of two quotients taken as one division */
void pgo_grads(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr Jinv, pgo_arr grads, pgo_box r) {
  const int ngrad = c->ne - 1;
  FOR_BOX(r) {
    const int off[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double rp[3], rm[3], inv[3];
    for (int a = 0; a < 3; a++) {
      rp[a] = *at(Q, i + off[a][0], j + off[a][1], k + off[a][2], 0);
      rm[a] = *at(Q, i - off[a][0], j - off[a][1], k - off[a][2], 0);
      inv[a] = 1.0 / (rp[a] * rm[a]);
    }
    double J[3][3];
    for (int a = 0; a < 3; a++)
      for (int e = 0; e < 3; e++) J[a][e] = *at(Jinv, i, j, k, 3 * a + e);
    for (int l = 0; l < ngrad; l++) {
      double d[3];
      for (int a = 0; a < 3; a++) {
        const int ip = i + off[a][0], jp = j + off[a][1], kp = k + off[a][2];
        const int im = i - off[a][0], jm = j - off[a][1], km = k - off[a][2];
        if (l == 3) {
          d[a] = 0.5 * (*at(state, ip, jp, kp, 1) - *at(state, im, jm, km, 1));
        } else {
          const int m = l + 1;
          d[a] = 0.5 * (*at(Q, ip, jp, kp, m) * rm[a] - *at(Q, im, jm, km, m) * rp[a]) * inv[a];
        }
      }
      for (int e = 0; e < 3; e++) *at(grads, i, j, k, 3 * l + e) = d[0] * J[0][e] + d[1] * J[1][e] + d[2] * J[2][e];
    }
  }
}

/* navstokes/kernels/flux/alpha-damping.mako:3-99 (Nishikawa's alpha damping): each face
 * gradient is the sides' mean damped toward the jump between the values each side extrapolates
 * to the face; Newtonian stress (lambda = -2 mu / 3), Fourier heat, Fick species fluxes with the
 * correction velocity, their enthalpy flux and the stress work, added to F */
void pgo_diffflux(const pgo_case *c, int axis, double alpha, pgo_arr Q, pgo_arr state, pgo_arr trstate,
                  pgo_arr hsp, pgo_arr grads, pgo_arr xcc, pgo_arr xf, pgo_arr S, pgo_arr F, pgo_box r) {
  const int ns = c->ns;
  const int dl[3] = {axis == 0, axis == 1, axis == 2};
  FOR_BOX(r) {
    const int iL = i - dl[0], jL = j - dl[1], kL = k - dl[2];
    const double bulkvisc = 0.0;
    const double Sx = *at(S, i, j, k, 0), Sy = *at(S, i, j, k, 1), Sz = *at(S, i, j, k, 2);
    const double Sabs = sqrt(Sx * Sx + Sy * Sy + Sz * Sz), Sinv = 1.0 / Sabs;
    const double nx = Sx * Sinv, ny = Sy * Sinv, nz = Sz * Sinv;
    const double mu = 0.5 * (*at(trstate, i, j, k, 0) + *at(trstate, iL, jL, kL, 0));
    const double kappa = 0.5 * (*at(trstate, i, j, k, 1) + *at(trstate, iL, jL, kL, 1));
    const double lambda = bulkvisc - 2.0 / 3.0 * mu;
    const double rhoinvL = *at(state, iL, jL, kL, 2), rhoinvR = *at(state, i, j, k, 2);
    const double uL = *at(Q, iL, jL, kL, 1) * rhoinvL, vL = *at(Q, iL, jL, kL, 2) * rhoinvL,
                 wL = *at(Q, iL, jL, kL, 3) * rhoinvL;
    const double uR = *at(Q, i, j, k, 1) * rhoinvR, vR = *at(Q, i, j, k, 2) * rhoinvR, wR = *at(Q, i, j, k, 3) * rhoinvR;
    const double ex = *at(xcc, i, j, k, 0) - *at(xcc, iL, jL, kL, 0);
    const double ey = *at(xcc, i, j, k, 1) - *at(xcc, iL, jL, kL, 1);
    const double ez = *at(xcc, i, j, k, 2) - *at(xcc, iL, jL, kL, 2);
    const double en0 = ex * nx, en1 = ey * ny, en2 = ez * nz;
    const double dampinv = alpha / sqrt(en0 * en0 + en1 * en1 + en2 * en2);
    const double dampx = dampinv * nx, dampy = dampinv * ny, dampz = dampinv * nz;
    const double xLx = *at(xf, i, j, k, 0) - *at(xcc, iL, jL, kL, 0), xLy = *at(xf, i, j, k, 1) - *at(xcc, iL, jL, kL, 1),
                 xLz = *at(xf, i, j, k, 2) - *at(xcc, iL, jL, kL, 2);
    const double xRx = *at(xf, i, j, k, 0) - *at(xcc, i, j, k, 0), xRy = *at(xf, i, j, k, 1) - *at(xcc, i, j, k, 1),
                 xRz = *at(xf, i, j, k, 2) - *at(xcc, i, j, k, 2);
#define FACEGRAD(l, phiL, phiR, gx, gy, gz)                                                                  \
  do {                                                                                                       \
    const double gLx = *at(grads, iL, jL, kL, 3 * (l)), gLy = *at(grads, iL, jL, kL, 3 * (l) + 1),           \
                 gLz = *at(grads, iL, jL, kL, 3 * (l) + 2);                                                  \
    const double gRx = *at(grads, i, j, k, 3 * (l)), gRy = *at(grads, i, j, k, 3 * (l) + 1),                 \
                 gRz = *at(grads, i, j, k, 3 * (l) + 2);                                                     \
    const double wl = (phiL) + gLx * xLx + gLy * xLy + gLz * xLz;                                            \
    const double wr = (phiR) + gRx * xRx + gRy * xRy + gRz * xRz;                                            \
    const double jump = wr - wl;                                                                             \
    gx = 0.5 * (gRx + gLx) + dampx * jump;                                                                   \
    gy = 0.5 * (gRy + gLy) + dampy * jump;                                                                   \
    gz = 0.5 * (gRz + gLz) + dampz * jump;                                                                   \
  } while (0)
    double dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz;
    FACEGRAD(0, uL, uR, dudx, dudy, dudz);
    FACEGRAD(1, vL, vR, dvdx, dvdy, dvdz);
    FACEGRAD(2, wL, wR, dwdx, dwdy, dwdz);
    const double div = dudx + dvdy + dwdz;
    const double txx = -2.0 * mu * dudx - lambda * div;
    const double txy = -mu * (dvdx + dudy);
    const double txz = -mu * (dwdx + dudz);
    const double tyy = -2.0 * mu * dvdy - lambda * div;
    const double tyz = -mu * (dwdy + dvdz);
    const double tzz = -2.0 * mu * dwdz - lambda * div;
    *at(F, i, j, k, 1) += txx * Sx + txy * Sy + txz * Sz;
    *at(F, i, j, k, 2) += txy * Sx + tyy * Sy + tyz * Sz;
    *at(F, i, j, k, 3) += txz * Sx + tyz * Sy + tzz * Sz;
    double dTx, dTy, dTz;
    FACEGRAD(3, *at(state, iL, jL, kL, 1), *at(state, i, j, k, 1), dTx, dTy, dTz);
    const double heatflux = -kappa * (dTx * Sx + dTy * Sy + dTz * Sz);
    const double uf = 0.5 * (uR + uL), vf = 0.5 * (vR + vL), wf = 0.5 * (wR + wL);
    double f4 = (uf * txx + vf * txy + wf * txz) * Sx + (uf * txy + vf * tyy + wf * tyz) * Sy +
                (uf * txz + vf * tyz + wf * tzz) * Sz + heatflux;
    double YL[MAXNS], YR[MAXNS], YlastL = 1.0, YlastR = 1.0;
    for (int n = 0; n < ns - 1; n++) {
      YL[n] = *at(Q, iL, jL, kL, 5 + n) * rhoinvL;
      YR[n] = *at(Q, i, j, k, 5 + n) * rhoinvR;
      YlastL -= *at(Q, iL, jL, kL, 5 + n) * rhoinvL;
      YlastR -= *at(Q, i, j, k, 5 + n) * rhoinvR;
    }
    (void)YlastL;
    (void)YlastR;
    double Dk, Vc = 0.0, gradYns = 0.0, sumYh = 0.0, Yns = 1.0;
    const double rho = 0.5 * (*at(Q, i, j, k, 0) + *at(Q, iL, jL, kL, 0));
    for (int n = 0; n < ns - 1; n++) {
      Dk = 0.5 * (*at(trstate, i, j, k, 2 + n) + *at(trstate, iL, jL, kL, 2 + n));
      double dYx, dYy, dYz;
      FACEGRAD(4 + n, YL[n], YR[n], dYx, dYy, dYz);
      const double gradYk = dYx * Sx + dYy * Sy + dYz * Sz;
      gradYns -= gradYk;
      Vc += Dk * gradYk;
      const double Jk = -rho * Dk * gradYk;
      const double hk = 0.5 * (*at(hsp, i, j, k, n) + *at(hsp, iL, jL, kL, n));
      const double Yk = 0.5 * (YR[n] + YL[n]);
      Yns -= Yk;
      sumYh += Yk * hk;
      *at(F, i, j, k, 5 + n) += Jk;
      f4 += Jk * hk;
    }
#undef FACEGRAD
    Dk = 0.5 * (*at(trstate, i, j, k, 2 + ns - 1) + *at(trstate, iL, jL, kL, 2 + ns - 1));
    Vc += Dk * gradYns;
    const double corr = rho * Vc;
    for (int n = 0; n < ns - 1; n++) *at(F, i, j, k, 5 + n) += 0.5 * (YR[n] + YL[n]) * corr;
    f4 += corr * sumYh;
    const double hns = 0.5 * (*at(hsp, i, j, k, ns - 1) + *at(hsp, iL, jL, kL, ns - 1));
    f4 += (-rho * Dk * gradYns + Yns * corr) * hns;
    *at(F, i, j, k, 4) += f4;
  }
}

/* ------------------------------------------------------------------------
 * the adiabatic slip wall: euler/kernels/bcs/slp-adia-wall.mako, bcs/wall.mako;
 * navstokes/kernels/bcs/slp-adia-wall.mako, bcs/wall.mako
 * ---------------------------------------------------------------------- */

/* solution.mako:14-23 fromprims for one cell */
static void fromprims_cell(const pgo_case *c, double p, double T, const double u[3], const double *Y, pgo_arr Q,
                           pgo_arr state, int i, int j, int k) {
  mixmom m;
  mix(c, Y, &m);
  eosstate s;
  frompT(c, &m, Y, p, T, &s);
  *at(Q, i, j, k, 0) = s.rho;
  *at(Q, i, j, k, 1) = s.rho * u[0];
  *at(Q, i, j, k, 2) = s.rho * u[1];
  *at(Q, i, j, k, 3) = s.rho * u[2];
  *at(Q, i, j, k, 4) = s.rho * (s.e + 0.5 * ((u[0]) * (u[0]) + (u[1]) * (u[1]) + (u[2]) * (u[2])));
  for (int n = 0; n < c->ns - 1; n++) *at(Q, i, j, k, 5 + n) = s.rho * Y[n];
  store(state, i, j, k, &s);
  storeZ(c, state, i, j, k, &s);
}

/* the walls (simulator/euler/bcs.py:15-27, navstokes/bcs.py:27-53): each halo cell h planes out
 * from the interior cell h planes in. kind 0 slp-adia (euler/kernels/bcs/slp-adia-wall.mako: the
 * momentum's normal part taken off twice, the state copied), 1 slp-isot (slp-isot-wall.mako: the
 * velocity's normal part taken off twice, the image's p and Y at T_w by the EOS), 2 no-slp-adia and
 * 4 mov-adia (navstokes wall.mako stuckhalo: the velocity reflected about the wall's, the energy
 * carrying the kinetic change, the state copied), 3 no-slp-isot and 5 mov-isot (stuckhaloisot) */
void pgo_wall_mirror(const pgo_case *c, int axis, int side, int kind, const double *wall, double Tw, const int *ncell,
                     int ng, int nstate, pgo_arr Q, pgo_arr state, pgo_arr S) {
  const int n = ncell[axis];
  const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
  for (int h = 0; h < ng; h++) {
    /* a block thinner than the halo reflects its last cell into every deeper layer
     * (launch/policy.py:168-184) */
    const int hm = h < n ? h : n - 1;
    const int halo = side ? ng + n + h : ng - 1 - h, mirror = side ? ng + n - 1 - hm : ng + hm, face = side ? ng + n : ng;
#pragma omp parallel for collapse(2) if ((int64_t)ncell[t1] * ncell[t2] > PGO_PAR_CELLS)
    for (int b2 = ng; b2 < ng + ncell[t2]; b2++)
      for (int b1 = ng; b1 < ng + ncell[t1]; b1++) {
        int ih[3], im[3], f[3];
        ih[axis] = halo, im[axis] = mirror, f[axis] = face;
        ih[t1] = im[t1] = f[t1] = b1;
        ih[t2] = im[t2] = f[t2] = b2;
        const double Sx = *at(S, f[0], f[1], f[2], 0), Sy = *at(S, f[0], f[1], f[2], 1), Sz = *at(S, f[0], f[1], f[2], 2);
        const double Sinv = 1.0 / sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
        const double nx = Sx * Sinv, ny = Sy * Sinv, nz = Sz * Sinv;
        if (kind == 0) {
          const double q1 = *at(Q, im[0], im[1], im[2], 1), q2 = *at(Q, im[0], im[1], im[2], 2),
                       q3 = *at(Q, im[0], im[1], im[2], 3);
          const double mn = 2.0 * (q1 * nx + q2 * ny + q3 * nz);
          *at(Q, ih[0], ih[1], ih[2], 0) = *at(Q, im[0], im[1], im[2], 0);
          *at(Q, ih[0], ih[1], ih[2], 1) = q1 - mn * nx;
          *at(Q, ih[0], ih[1], ih[2], 2) = q2 - mn * ny;
          *at(Q, ih[0], ih[1], ih[2], 3) = q3 - mn * nz;
          for (int l = 4; l < c->ne; l++) *at(Q, ih[0], ih[1], ih[2], l) = *at(Q, im[0], im[1], im[2], l);
          for (int f8 = 0; f8 < nstate; f8++) *at(state, ih[0], ih[1], ih[2], f8) = *at(state, im[0], im[1], im[2], f8);
        } else if (kind == 2 || kind == 4) {
          const double rho = *at(Q, im[0], im[1], im[2], 0), rhoinv = 1.0 / rho;
          double ke = 0.0, kehalo = 0.0;
          for (int d = 0; d < 3; d++) {
            const double u = *at(Q, im[0], im[1], im[2], 1 + d) * rhoinv, uh = 2.0 * wall[d] - u;
            *at(Q, ih[0], ih[1], ih[2], 1 + d) = rho * uh;
            ke += u * u;
            kehalo += uh * uh;
          }
          *at(Q, ih[0], ih[1], ih[2], 0) = rho;
          *at(Q, ih[0], ih[1], ih[2], 4) = *at(Q, im[0], im[1], im[2], 4) + 0.5 * rho * (kehalo - ke);
          for (int l = 5; l < c->ne; l++) *at(Q, ih[0], ih[1], ih[2], l) = *at(Q, im[0], im[1], im[2], l);
          for (int f8 = 0; f8 < nstate; f8++) *at(state, ih[0], ih[1], ih[2], f8) = *at(state, im[0], im[1], im[2], f8);
        } else {
          const double rhoinv = 1.0 / *at(Q, im[0], im[1], im[2], 0);
          double u[3], Y[MAXNS];
          if (kind == 1) {
            const double uin = *at(Q, im[0], im[1], im[2], 1) * rhoinv, vin = *at(Q, im[0], im[1], im[2], 2) * rhoinv,
                         win = *at(Q, im[0], im[1], im[2], 3) * rhoinv;
            const double un = 2.0 * (uin * nx + vin * ny + win * nz);
            u[0] = uin - un * nx;
            u[1] = vin - un * ny;
            u[2] = win - un * nz;
          } else {
            for (int d = 0; d < 3; d++) u[d] = 2.0 * wall[d] - *at(Q, im[0], im[1], im[2], 1 + d) * rhoinv;
          }
          for (int l = 0; l < c->ns - 1; l++) Y[l] = *at(Q, im[0], im[1], im[2], 5 + l) * rhoinv;
          fromprims_cell(c, *at(state, im[0], im[1], im[2], 0), Tw, u, Y, Q, state, ih[0], ih[1], ih[2]);
        }
      }
  }
}

/* the wall's flux from the cell next to it, written over F: p_w = max(p + rho c u_n, 0)
 * (euler/kernels/bcs/wall.mako wallpressure: the acoustic star pressure to first order, u_n toward
 * the wall); a viscous wall adds the stress of the two-point normal velocity gradient relative to
 * the wall, every component where the flow is held (stuck) and the normal one where it slips, the
 * wall's work, and for an isothermal wall the heat flux from T_w with mu and kappa at the wall
 * (navstokes/kernels/bcs/wall.mako wallfaceflux, adiaflux, isotflux, wallcoeffs) */
void pgo_wall_flux(const pgo_case *c, int axis, int side, int kind, const double *wall, double Tw, const int *ncell,
                   int ng, int viscous, pgo_arr Q, pgo_arr state, pgo_arr trstate, pgo_arr xcc, pgo_arr xf, pgo_arr S,
                   pgo_arr F) {
  const int n = ncell[axis];
  const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
  const int face = side ? ng + n : ng, cell = side ? ng + n - 1 : ng;
  const double outward = side ? 1.0 : -1.0;
  const int isot = kind == 1 || kind == 3 || kind == 5;
  const double stuck = kind >= 2 ? 1.0 : 0.0;
#pragma omp parallel for collapse(2) if ((int64_t)ncell[t1] * ncell[t2] > PGO_PAR_CELLS)
  for (int b2 = ng; b2 < ng + ncell[t2]; b2++)
    for (int b1 = ng; b1 < ng + ncell[t1]; b1++) {
      int f[3], ic[3];
      f[axis] = face, ic[axis] = cell;
      f[t1] = ic[t1] = b1;
      f[t2] = ic[t2] = b2;
      const double S3[3] = {*at(S, f[0], f[1], f[2], 0), *at(S, f[0], f[1], f[2], 1), *at(S, f[0], f[1], f[2], 2)};
      const double q[4] = {*at(Q, ic[0], ic[1], ic[2], 0), *at(Q, ic[0], ic[1], ic[2], 1), *at(Q, ic[0], ic[1], ic[2], 2),
                           *at(Q, ic[0], ic[1], ic[2], 3)};
      const double area = sqrt(S3[0] * S3[0] + S3[1] * S3[1] + S3[2] * S3[2]);
      const double un = outward * (q[1] * S3[0] + q[2] * S3[1] + q[3] * S3[2]) / (q[0] * area);
      const double pw = fmax(*at(state, ic[0], ic[1], ic[2], 0) + q[0] * *at(state, ic[0], ic[1], ic[2], 7) * un, 0.0);
      double Fv[MAXNS + 4];
      Fv[0] = 0.0;
      Fv[1] = pw * S3[0];
      Fv[2] = pw * S3[1];
      Fv[3] = pw * S3[2];
      for (int l = 4; l < c->ne; l++) Fv[l] = 0.0;
      if (viscous) {
        double mu, kappa, dT;
        if (isot) {
          const double rhoinv = 1.0 / q[0];
          double Y[MAXNS], Ylast = 1.0;
          for (int l = 0; l < c->ns - 1; l++) {
            Y[l] = *at(Q, ic[0], ic[1], ic[2], 5 + l) * rhoinv;
            Ylast -= *at(Q, ic[0], ic[1], ic[2], 5 + l) * rhoinv;
          }
          Y[c->ns - 1] = fmax(Ylast, 0.0);
          mixmom m;
          mix(c, Y, &m);
          eosstate thermo;
          frompT(c, &m, Y, *at(state, ic[0], ic[1], ic[2], 0), Tw, &thermo);
          double denom = 0.0, X[MAXNS];
          for (int l = 0; l < c->ns; l++) denom += Y[l] * c->MWinv[l];
          const double MWmix = 1.0 / denom;
          for (int l = 0; l < c->ns; l++) X[l] = Y[l] * c->MWinv[l] * MWmix;
          mu = mu_kappa(c, X, Tw, ELOG(Tw), thermo.rho, &kappa);
          dT = Tw - *at(state, ic[0], ic[1], ic[2], 1);
        } else {
          mu = *at(trstate, ic[0], ic[1], ic[2], 0);
          kappa = *at(trstate, ic[0], ic[1], ic[2], 1);
          dT = 0.0;
        }
        const double Sabs = sqrt(S3[0] * S3[0] + S3[1] * S3[1] + S3[2] * S3[2]), Sinv = 1.0 / Sabs;
        const double nv[3] = {S3[0] * Sinv, S3[1] * Sinv, S3[2] * Sinv};
        const double dn = fabs((*at(xcc, ic[0], ic[1], ic[2], 0) - *at(xf, f[0], f[1], f[2], 0)) * nv[0] +
                               (*at(xcc, ic[0], ic[1], ic[2], 1) - *at(xf, f[0], f[1], f[2], 1)) * nv[1] +
                               (*at(xcc, ic[0], ic[1], ic[2], 2) - *at(xf, f[0], f[1], f[2], 2)) * nv[2]);
        const double rhoinv = 1.0 / q[0];
        double dun[3];
        for (int d = 0; d < 3; d++) dun[d] = outward * (wall[d] - q[1 + d] * rhoinv) / dn;
        const double div = dun[0] * nv[0] + dun[1] * nv[1] + dun[2] * nv[2];
        for (int d = 0; d < 3; d++) dun[d] = div * nv[d] + stuck * (dun[d] - div * nv[d]);
        double work = 0.0;
        for (int d = 0; d < 3; d++) {
          const double tS = -mu * Sabs * (dun[d] + div * nv[d] / 3.0);
          Fv[1 + d] += tS;
          work += wall[d] * tS;
        }
        Fv[4] += work - kappa * outward * (dT) / dn * Sabs;
      }
      for (int l = 0; l < c->ne; l++) *at(F, f[0], f[1], f[2], l) = Fv[l];
    }
}

static pgo_box interior(const pgo_block *b) {
  pgo_box r;
  for (int a = 0; a < 3; a++) {
    r.start[a] = b->ng;
    r.extent[a] = b->ncell[a];
  }
  return r;
}

static pgo_box facebox(const pgo_block *b, int axis) {
  pgo_box r = interior(b);
  r.extent[axis] += 1;
  return r;
}

/* the halo slab of one interface, depth planes deep, over the interior of the other axes */
static pgo_box haloslab(const pgo_block *b, int a, int side, int depth) {
  pgo_box r = interior(b);
  r.start[a] = side ? b->ng + b->ncell[a] : b->ng - depth;
  r.extent[a] = depth;
  return r;
}

/* simulator/advec/simulator.py:57-75: the exchange and the halo EOS on periodic faces, the
 * walls' mirror halos (preadvflux), the source; for a viscous case the transport over the
 * interior and the first halo layer of each interface ("shallow", launch/ranges.py:52-53;
 * advecdiff/simulator.py postbc) */
static void prerhs(const pgo_case *c, pgo_block *b, pgo_arr Q, pgo_arr dQ, int withdQ) {
  pgo_arr S[3] = {b->iS, b->jS, b->kS};
  for (int a = 0; a < 3; a++)
    if (!b->bc[2 * a]) pgo_fill_periodic(Q, c->ne, b->ncell, b->ng, b->ng, a);
  for (int a = 0; a < 3; a++)
    for (int side = 0; side < 2; side++)
      if (!b->bc[2 * a + side]) pgo_fromQ(c, Q, b->state, haloslab(b, a, side, b->ng));
  for (int a = 0; a < 3; a++)
    for (int side = 0; side < 2; side++)
      if (b->bc[2 * a + side])
        pgo_wall_mirror(c, a, side, b->bc[2 * a + side] - 1, b->bcwall[2 * a + side], b->bcT[2 * a + side], b->ncell,
                        b->ng, b->nstate, Q, b->state, S[a]);
  if (withdQ && b->chemistry == 1) pgo_source(c, Q, b->state, dQ, interior(b));
  if (withdQ && b->chemistry == 2)
    pgo_substep(c, Q, b->state, dQ, 1, b->substeps, b->subcap, b->subeqtol, b->dt, interior(b));
  if (withdQ && b->chemistry == 3)
    pgo_implicit(c, Q, b->state, dQ, 1, b->imparts, b->dt, interior(b));
  if (b->viscous) {
    pgo_transport(c, Q, b->state, b->trstate, b->hsp, interior(b));
    for (int a = 0; a < 3; a++)
      for (int side = 0; side < 2; side++)
        if (!b->bc[2 * a + side]) pgo_transport(c, Q, b->state, b->trstate, b->hsp, haloslab(b, a, side, 1));
    /* the sponge over the cells the transport has just made (viscsponge.py: the "shallow" range) */
    if (b->sponge) {
      pgo_viscsponge(b->xcc, b->trstate, b->spongeline, interior(b));
      for (int a = 0; a < 3; a++)
        for (int side = 0; side < 2; side++)
          if (!b->bc[2 * a + side]) pgo_viscsponge(b->xcc, b->trstate, b->spongeline, haloslab(b, a, side, 1));
    }
  }
}

/* advec/simulator.py:76-80; advecdiff/simulator.py:20-35: gradients over the interior and their
 * one-plane exchange (a wall's halo gradients are read by nothing), the advective flux, the
 * viscous flux added to it, the walls' own flux through their faces (postflux), the divergence */
static void rhs(const pgo_case *c, pgo_block *b, pgo_arr Q, pgo_arr dQ) {
  pgo_arr F[3] = {b->iF, b->jF, b->kF}, S[3] = {b->iS, b->jS, b->kS};
  if (b->viscous) {
    pgo_grads(c, Q, b->state, b->Jinv, b->grads, interior(b));
    for (int a = 0; a < 3; a++)
      if (!b->bc[2 * a]) pgo_fill_periodic(b->grads, 3 * (c->ne - 1), b->ncell, b->ng, 1, a);
  }
  for (int a = 0; a < 3; a++) {
    pgo_flux(c, a, Q, b->state, S[a], F[a], facebox(b, a));
    if (b->viscous)
      pgo_diffflux(c, a, b->alpha, Q, b->state, b->trstate, b->hsp, b->grads, b->xcc, b->xf[a], S[a], F[a],
                   facebox(b, a));
  }
  for (int a = 0; a < 3; a++)
    for (int side = 0; side < 2; side++)
      if (b->bc[2 * a + side])
        pgo_wall_flux(c, a, side, b->bc[2 * a + side] - 1, b->bcwall[2 * a + side], b->bcT[2 * a + side], b->ncell,
                      b->ng, b->viscous, Q, b->state, b->trstate, b->xcc, b->xf[a], S[a], F[a]);
  pgo_applyflux(c->ne, dQ, b->iF, b->jF, b->kF, b->detJinv, b->chemistry, interior(b));
}

/* the registers a stage reads and writes, and its combine weights */
void pgo_advance(const pgo_case *c, pgo_block *b, int n, double dt) {
  const pgo_box r = interior(b);
  pgo_arr start = b->regs[0], stage = b->regs[1], dQ = b->regs[b->nregs - 1];
  if (b->integrator == 3) {
    pgo_arr total = b->regs[2];
    static const double wdQ[4] = {0.5, 0.5, 1.0, -1.0}, wsum[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double w1[1] = {wsum[n]};
    pgo_combine(c->ne, total, n != 0, 1.0, 1, &dQ, w1, dt, 1, r);
    if (n == 3) {
      const double w2[1] = {1.0};
      pgo_combine(c->ne, start, 1, 1.0, 1, &total, w2, dt, 0, r);
    } else {
      pgo_arr X[2] = {start, dQ};
      const double w2[2] = {1.0, wdQ[n]};
      pgo_combine(c->ne, stage, 0, 0.0, 2, X, w2, dt, 1, r);
    }
    return;
  }
  /* (wQ0, wQ, wdQ) per stage; Qout = wQ Q + wQ0 Qstart + wdQ dt dQ */
  static const double rk1[1][3] = {{0.0, 1.0, 1.0}};
  static const double rk2[2][3] = {{0.0, 1.0, 1.0}, {0.5, 0.5, 0.5}};
  static const double rk3[3][3] = {{0.0, 1.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  const double *wt = b->integrator == 0 ? rk1[n] : b->integrator == 1 ? rk2[n] : rk3[n];
  const int nst = b->integrator + 1, last = nst - 1;
  pgo_arr Q = n == 0 ? start : stage, Qout = n == last ? start : stage;
  const int Qis0 = n == 0, outis0 = n == last;
  /* the weights per distinct register, Q first, then the start, then dQ */
  double wQ = wt[1], w0 = wt[0];
  if (Qis0) {
    wQ += w0;
    w0 = 0.0;
  }
  double own = 0.0;
  pgo_arr X[3];
  double w[3];
  int nt = 0;
  /* Q's weight */
  if ((Qis0 && outis0) || (!Qis0 && !outis0)) {
    own = wQ;
  } else if (wQ != 0.0) {
    X[nt] = Q;
    w[nt++] = wQ;
  }
  if (!Qis0) {
    if (outis0)
      own = w0;
    else if (w0 != 0.0) {
      X[nt] = start;
      w[nt++] = w0;
    }
  }
  X[nt] = dQ;
  w[nt++] = wt[2];
  pgo_combine(c->ne, Qout, own != 0.0, own, nt, X, w, dt, 1, r);
}

void pgo_stage(const pgo_case *c, pgo_block *b, int n, double dt) {
  const int nst = b->integrator == 3 ? 4 : b->integrator + 1, last = nst - 1;
  pgo_arr Q = n == 0 ? b->regs[0] : b->regs[1], dQ = b->regs[b->nregs - 1];
  pgo_arr Qout = n == last ? b->regs[0] : b->regs[1];
  b->dt = dt;
  prerhs(c, b, Q, dQ, 1);
  rhs(c, b, Q, dQ);
  pgo_advance(c, b, n, dt);
  pgo_fromQ(c, Qout, b->state, interior(b));
}

void pgo_rhs(const pgo_case *c, pgo_block *b, int q, int dq) {
  prerhs(c, b, b->regs[q], b->regs[dq], 1);
  rhs(c, b, b->regs[q], b->regs[dq]);
}

void pgo_step(const pgo_case *c, pgo_block *b, double dt) {
  const int nst = b->integrator == 3 ? 4 : b->integrator + 1;
  for (int n = 0; n < nst; n++) pgo_stage(c, b, n, dt);
}

/* euler/kernels/cflmax.mako:3-33: per interior cell the acoustic speed c sqrt(sum 1/d^2) over
 * the acoustic length, the convective sum of |u . n| / d per axis (n the mean of the cell's two
 * faces' unit normals) and their sum; folded by fmax from -inf (codegen/grammar.py:34-42) */
void pgo_cflmax(pgo_arr Q, pgo_arr state, pgo_arr dijk, pgo_arr iS, pgo_arr jS, pgo_arr kS, pgo_box r, double out[3]) {
  const pgo_arr S[3] = {iS, jS, kS};
  double best[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int k = r.start[2]; k < r.start[2] + r.extent[2]; k++)
    for (int j = r.start[1]; j < r.start[1] + r.extent[1]; j++)
      for (int i = r.start[0]; i < r.start[0] + r.extent[0]; i++) {
        const double di = *at(dijk, i, j, k, 0), dj = *at(dijk, i, j, k, 1), dk = *at(dijk, i, j, k, 2);
        const double rhoinv = 1.0 / *at(Q, i, j, k, 0);
        const double u = *at(Q, i, j, k, 1) * rhoinv, v = *at(Q, i, j, k, 2) * rhoinv, w = *at(Q, i, j, k, 3) * rhoinv;
        double un[3];
        for (int a = 0; a < 3; a++) {
          const int hi[3] = {i + (a == 0), j + (a == 1), k + (a == 2)};
          const double L0 = *at(S[a], i, j, k, 0), L1 = *at(S[a], i, j, k, 1), L2 = *at(S[a], i, j, k, 2);
          const double H0 = *at(S[a], hi[0], hi[1], hi[2], 0), H1 = *at(S[a], hi[0], hi[1], hi[2], 1),
                       H2 = *at(S[a], hi[0], hi[1], hi[2], 2);
          const double S0inv = 1.0 / sqrt(L0 * L0 + L1 * L1 + L2 * L2);
          const double S1inv = 1.0 / sqrt(H0 * H0 + H1 * H1 + H2 * H2);
          const double nx = 0.5 * (L0 * S0inv + H0 * S1inv);
          const double ny = 0.5 * (L1 * S0inv + H1 * S1inv);
          const double nz = 0.5 * (L2 * S0inv + H2 * S1inv);
          un[a] = fabs(nx * u + ny * v + nz * w);
        }
        const double c = *at(state, i, j, k, 7);
        const double acoustic = c * sqrt(1.0 / (di * di) + 1.0 / (dj * dj) + 1.0 / (dk * dk));
        const double convective = un[0] / di + un[1] / dj + un[2] / dk;
        best[0] = fmax(best[0], acoustic);
        best[1] = fmax(best[1], convective);
        best[2] = fmax(best[2], acoustic + convective);
      }
  for (int n = 0; n < 3; n++) out[n] = best[n];
}

/* chemint/kernels/substep.mako:5-93 with bulk.mako's one-pass form (bulk-substeps = max-substeps):
 * each cell of the box integrates its chemistry over dt in substeps as long as its limiting species
 * allows, the equilibrium exit judged at the first substep when eqtol > 0, and writes the mean rate
 * over the step (into dQ rows 5.. with rows 0..4 zero when begin_dQ, else into rows 0..ns-2 of out);
 * substeps[box index, i fastest] receives each cell's count */
void pgo_substep(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr out, int begin_dQ, int32_t *substeps, int cap,
                 double eqtol, double dt, pgo_box r) {
  const int ns = c->ns, nt = c->nt;
  const double eps = DBL_EPSILON, tiny = DBL_MIN, stepfrac = 1e-12;
  FOR_BOX(r) {
    const double rho = *at(Q, i, j, k, 0), rhoinv = 1.0 / rho, step = dt, stepinv = 1.0 / step;
    double Y[MAXNS], src[MAXNS], rate[MAXNS], D[MAXNS];
    /* massfrac (solution.mako:25-35) */
    double Ylast = 1.0, Ysum = 0.0;
    for (int n = 0; n < ns - 1; n++) {
      const double Yn = fmax(fmin(*at(Q, i, j, k, 5 + n) * rhoinv, 1.0), 0.0);
      Y[n] = Yn;
      Ylast -= Yn;
      Ysum += Yn;
    }
    const double Yscale = 1.0 / fmax(Ysum, 1.0);
    for (int n = 0; n < ns - 1; n++) Y[n] *= Yscale;
    Y[ns - 1] = fmax(Ylast, 0.0);
    for (int n = 0; n < ns; n++) src[n] = 0.0;
    double T = *at(state, i, j, k, 1), remaining = 1.0;
    const int exitq = eqtol > 0.0;
    int sub = 0, early = 0;
    for (; sub < cap && remaining > stepfrac; sub++) {
      pgo_sources(c, rho, Y, T, rate, exitq && sub == 0 ? D : NULL);
      double abest = 0.0, bbest = 0.0;
      for (int n = 0; n < ns; n++) {
        const double sign = copysign(1.0, rate[n]);
        const double headroom = 0.5 * ((1.0 - sign) * Y[n] + (1.0 + sign) * (1.0 - Y[n]));
        const double a = rho * fmax(headroom, eps), b = fmax(fabs(rate[n]), tiny);
        if (n == 0 || a * bbest < abest * b) {
          abest = a;
          bbest = b;
        }
      }
      const double share = fmin(remaining, abest / bbest * stepinv);
      /* the substep reads ln T from the rate state's table log (finiterate.mako:143) */
      const double u = pgo_tlog(T), Tinv = 1.0 / T;
      double cv = 0.0, dTdt = 0.0, e[MAXNS];
      for (int n = 0; n < ns; n++) {
        double cpR = 0.0, hRT = 0.0;
        for (int m = nt - 1; m >= 0; m--) cpR = cpR * u + c->cppoly[n * nt + m];
        for (int m = nt - 1; m >= 0; m--) hRT = hRT * u + c->hpoly[n * nt + m];
        hRT += c->h298[n] * Tinv;
        e[n] = (hRT - 1.0) * T * c->Rsp[n];
        cv += (cpR - 1.0) * c->Rsp[n] * Y[n];
        dTdt -= e[n] * rate[n];
      }
      const double cvrhoinv = 1.0 / (cv * rho);
      dTdt *= cvrhoinv;
      if (exitq) {
        const double tleft = remaining * step, tleftinv = 1.0 / tleft;
        int settled = sub == 0 && tleft > 0.0 && fabs(dTdt) * tleft < 0.1;
        for (int n = 0; n < ns && settled; n++) settled = fabs(rate[n] * rhoinv) * tleft < eqtol;
        if (settled) {
          double dT = 0.0;
          for (int n = 0; n < ns; n++) {
            const double rhoden = rho * fmax(Y[n], c->cfloor * c->MW[n] * rhoinv);
            const double move = fmax((rate[n] * rhoinv) * tleft * rhoden / (rhoden + D[n] * tleft), -Y[n]);
            const double rr = move * rho * tleftinv;
            dT -= e[n] * rr;
            Y[n] += move;
            src[n] += rr * remaining;
          }
          T += dT * cvrhoinv * tleft;
          remaining = 0.0;
          early = 1;
          break;
        }
      }
      const double tsub = share * step;
      for (int n = 0; n < ns; n++) {
        Y[n] += (rate[n] * rhoinv) * tsub;
        src[n] += rate[n] * share;
      }
      T += dTdt * tsub;
      remaining -= share;
      if (share <= stepfrac) {
        remaining = 0.0;
        early = 1;
        break;
      }
    }
    const int64_t code = (i - r.start[0]) + (int64_t)r.extent[0] * ((j - r.start[1]) + (int64_t)r.extent[1] * (k - r.start[2]));
    substeps[code] = early ? sub + 1 : sub;
    if (begin_dQ) {
      for (int l = 0; l < 5; l++) *at(out, i, j, k, l) = 0.0;
      for (int n = 0; n < ns - 1; n++) *at(out, i, j, k, 5 + n) = src[n];
    } else {
      for (int n = 0; n < ns - 1; n++) *at(out, i, j, k, n) = src[n];
    }
  }
}

/* Implicit chemistry (no counterpart in the reference; the stiff alternative to its sub-stepped
 * form, chemint/substepped.py): the 0D system the sub-stepped form integrates, y = (Y_0 .. Y_{ns-1},
 * T) at constant density and internal energy with f = (w_n / rho, -sum e_n w_n / (rho cv))
 * (chemint/kernels/substep.mako:14-50), stepped over dt by the linearly implicit Euler method in
 * `parts` equal parts: y1 = y0 + h x with (I - h J) x = f(y0), J by forward differences at y0 rounded
 * to FP32 (the matrix of a state within its last bits is the same), the system solved by an LU in FP64
 * without pivoting. The mean rate rho (Y_end - Y_start) / dt is written as
 * pgo_substep writes its own. Every operation is FP64: this is the plain form the device's
 * FP32-factored, FP64-refined solve is held to. */
static void im_rhs0d(const pgo_case *c, double rho, double rhoinv, const double *y, double *f) {
  const int ns = c->ns, nt = c->nt;
  double rate[MAXNS];
  const double T = y[ns];
  pgo_sources(c, rho, y, T, rate, NULL);
  const double u = pgo_tlog(T), Tinv = 1.0 / T;
  double cv = 0.0, dTdt = 0.0;
  for (int n = 0; n < ns; n++) {
    double cpR = 0.0, hRT = 0.0;
    for (int m = nt - 1; m >= 0; m--) cpR = cpR * u + c->cppoly[n * nt + m];
    for (int m = nt - 1; m >= 0; m--) hRT = hRT * u + c->hpoly[n * nt + m];
    hRT += c->h298[n] * Tinv;
    const double e = (hRT - 1.0) * T * c->Rsp[n];
    cv += (cpR - 1.0) * c->Rsp[n] * y[n];
    dTdt -= e * rate[n];
  }
  const double cvrhoinv = 1.0 / (cv * rho);
  for (int n = 0; n < ns; n++) f[n] = rate[n] * rhoinv;
  f[ns] = dTdt * cvrhoinv;
}

/* the forward difference's step for unknown m: 2^-14 of T or of a mass fraction, at least 2^-14 * 1e-3
 * for a trace species (large on purpose: the rates' rounding over the step is what the difference adds
 * of its own, and a step of 2^-26 makes a stiff part's result move by 1e-6 with y's last bit), the
 * step taken being what the rounding leaves */
double pgo_fd_step(double ym, int temperature) {
  const double scale = temperature ? ym : fmax(fabs(ym), 1e-3);
  return (ym + 6.103515625e-05 * scale) - ym;
}

/* I - h J at y (nv = ns + 1 unknowns), row-major into A */
void pgo_implicit_matrix(const pgo_case *c, double rho, const double *y, double h, double *A) {
  const int ns = c->ns, nv = ns + 1;
  const double rhoinv = 1.0 / rho;
  double f0[MAXNS + 1], f1[MAXNS + 1], yp[MAXNS + 1];
  im_rhs0d(c, rho, rhoinv, y, f0);
  for (int m = 0; m < nv; m++) {
    const double dlt = pgo_fd_step(y[m], m == ns);
    for (int r = 0; r < nv; r++) yp[r] = r == m ? y[m] + dlt : y[r];
    im_rhs0d(c, rho, rhoinv, yp, f1);
    const double hd = h / dlt;
    for (int r = 0; r < nv; r++) A[r * nv + m] = (r == m ? 1.0 : 0.0) - (f1[r] - f0[r]) * hd;
  }
}

/* f of the 0D system at y, for the tests */
void pgo_implicit_rhs(const pgo_case *c, double rho, const double *y, double *f) { im_rhs0d(c, rho, 1.0 / rho, y, f); }

/* the right-looking LU without pivoting, in place (L unit lower below the diagonal) */
void pgo_lu_nopivot(int n, double *A) {
  for (int p = 0; p < n; p++) {
    const double rp = 1.0 / A[p * n + p];
    for (int i = p + 1; i < n; i++) {
      const double li = A[i * n + p] * rp;
      A[i * n + p] = li;
      for (int j = p + 1; j < n; j++) A[i * n + j] -= li * A[p * n + j];
    }
  }
}

void pgo_lu_apply(int n, const double *LU, double *z) {
  for (int r = 1; r < n; r++)
    for (int c = 0; c < r; c++) z[r] -= LU[r * n + c] * z[c];
  for (int r = n - 1; r >= 0; r--) {
    for (int c = r + 1; c < n; c++) z[r] -= LU[r * n + c] * z[c];
    z[r] = z[r] / LU[r * n + r];
  }
}

/* one cell: y (ns + 1 values) advanced in place over dt in `parts` linearly implicit Euler parts */
void pgo_implicit_cell(const pgo_case *c, double rho, double *y, double dt, int parts) {
  const int ns = c->ns, nv = ns + 1;
  const double h = dt / parts;
  double A[(MAXNS + 1) * (MAXNS + 1)], x[MAXNS + 1];
  for (int part = 0; part < parts; part++) {
    /* the matrix at y rounded to FP32 (a last-bit change of y then leaves it as it was) */
    double yr[MAXNS + 1];
    for (int r = 0; r < nv; r++) yr[r] = (double)(float)y[r];
    pgo_implicit_matrix(c, rho, yr, h, A);
    im_rhs0d(c, rho, 1.0 / rho, y, x);
    pgo_lu_nopivot(nv, A);
    pgo_lu_apply(nv, A, x);
    for (int r = 0; r < nv; r++) y[r] = y[r] + h * x[r];
  }
}

void pgo_implicit(const pgo_case *c, pgo_arr Q, pgo_arr state, pgo_arr out, int begin_dQ, int parts, double dt,
                  pgo_box r) {
  const int ns = c->ns;
  FOR_BOX(r) {
    const double rho = *at(Q, i, j, k, 0), rhoinv = 1.0 / rho;
    double y[MAXNS + 1], Ys[MAXNS];
    /* massfrac (solution.mako:25-35) */
    double Ylast = 1.0, Ysum = 0.0;
    for (int n = 0; n < ns - 1; n++) {
      const double Yn = fmax(fmin(*at(Q, i, j, k, 5 + n) * rhoinv, 1.0), 0.0);
      Ys[n] = Yn;
      Ylast -= Yn;
      Ysum += Yn;
    }
    const double Yscale = 1.0 / fmax(Ysum, 1.0);
    for (int n = 0; n < ns - 1; n++) Ys[n] *= Yscale;
    Ys[ns - 1] = fmax(Ylast, 0.0);
    for (int n = 0; n < ns; n++) y[n] = Ys[n];
    y[ns] = *at(state, i, j, k, 1);
    pgo_implicit_cell(c, rho, y, dt, parts);
    const double rdt = rho / dt;
    if (begin_dQ) {
      for (int l = 0; l < 5; l++) *at(out, i, j, k, l) = 0.0;
      for (int n = 0; n < ns - 1; n++) *at(out, i, j, k, 5 + n) = (y[n] - Ys[n]) * rdt;
    } else {
      for (int n = 0; n < ns - 1; n++) *at(out, i, j, k, n) = (y[n] - Ys[n]) * rdt;
    }
  }
}
