/*
 * moloch_oracle.c -- CPU FP64 restatement of the MOLOCH timestep.
 *
 * TEST INFRASTRUCTURE ONLY (see moloch_oracle.h).  This is the checker the
 * HIP implementation is compared against; it is never part of the product.
 *
 * Every routine follows nompi/mod_moloch.F90 of the reference (line ranges
 * cited per routine), keeping the reference's loop ranges, halo-fill
 * positions and left-to-right operand order.  It is built with
 * -ffp-contract=off so no fused multiply-add changes the rounding.
 *
 * Pinning status (what checks each routine; tests/test_oracle_*.py):
 *   constants, gzita, zita levels       -> reference modules built in
 *                                          oracle/_ref (golden vectors)
 *   pfesat/pfwsat                       -> Goff-Gratch closed form, clamps
 *   halo fills                          -> brute-force periodic index maps
 *   xknu, coru/corv                     -> closed forms
 *   thomas column                       -> dense solve of the assembled system
 *   divdamp                             -> Fourier symbol of the 5-point
 *                                          Laplacian, discrete gradient
 *   wafone                              -> constant/zero preservation, exact
 *                                          Courant-1 shift, textbook
 *                                          WAF-superbee (Toro) brute force
 *   stag<->cross interpolators          -> exactness on cubics
 *   prologue/epilogue                   -> potential temperature closed form,
 *                                          round trip, reference fill
 *                                          (mod_atmosphere.F90:210-222) is the
 *                                          inverse of the ps extrapolation
 *   whole step                          -> exact discrete rest state,
 *                                          translation equivariance (MPI halo
 *                                          order), mirror symmetry, seam
 *                                          locality of the nompi halo order
 *   acoustic substep phases             -> tests/test_oracle_physics_pins.py:
 *     Equation 10, surface w               kinematic condition and metric term
 *                                          in closed form (gzita golden)
 *     divergence (horizontal + s part)     uniform wind over terrain is
 *                                          divergence-free (truncation only)
 *     Exner update, pressure gradient      discrete Lamb-wave frequency of
 *                                          forward-backward stepping; terrain
 *                                          PGF balances the rest state
 *     gravity-wave term of zrom1w          anticipated interface theta from
 *                                          the physical level heights
 *     moist zdth branch                    latent heating of a lifted parcel
 *                                          (Clausius-Clapeyron vs es slope)
 *     fulleq coefficients                  moist heat capacities, to the
 *                                          reference's two figures
 *   tests/test_oracle_pin_mutants.py plants 18 plausible mistakes in a scratch
 *   copy of this file; each makes the pins fail.
 */
#include "moloch_oracle.h"

/* Loop-level threads for the CPU baseline build (-fopenmp, the library
 * libmoloch_oracle_omp.so).  Every marked loop nest writes only elements of
 * its own (j, i, k), or of its own column, and reads nothing the same nest
 * writes elsewhere, so the threaded build computes the same bits as the
 * serial one (tests/test_oracle_pins.py checks it).  Without -fopenmp the
 * marks are empty. */
#ifdef _OPENMP
#define ORA_PAR_IJ _Pragma("omp parallel for schedule(static) private(i, j)")
#define ORA_PAR_J _Pragma("omp parallel for schedule(static) private(j)")
#define ORA_PAR_COLUMNS _Pragma("omp parallel for schedule(static) private(j, k, zwexpl, zu, zd, wcol)")
#else
#define ORA_PAR_IJ
#define ORA_PAR_J
#define ORA_PAR_COLUMNS
#endif

#include <math.h>
#include <stddef.h>

/* ---------------------------------------------------------------- constants
 * mod_constants.F90: egrav :81, amd :103, amw :105, boltzk :88, navgdr :90,
 * rgasmol :126, rgas :130, rwat :131, cpd :141, cvd :145, wlhv :167,
 * regrav :178, rcpd :179, rovcp :180, rdrcv :181, cpovr :182, tzero :193,
 * p00 :225, lrate :231, mathpi :247, degrad :254, raddeg :255,
 * earthrad :275, eomeg :279, eomeg2 :280, ep1 :296, ep2 :299.
 */
#define K_EGRAV 9.80665
#define K_AMD 28.96454
#define K_AMW 18.01528
#define K_BOLTZK 1.3806490e-23
#define K_NAVGDR 6.02214076e23
static const double k_rgasmol = K_NAVGDR * K_BOLTZK;
static const double k_rgas = (K_NAVGDR * K_BOLTZK / K_AMD) * 1000.0;
static const double k_rwat = (K_NAVGDR * K_BOLTZK / K_AMW) * 1000.0;
#define RGAS ((K_NAVGDR * K_BOLTZK / K_AMD) * 1000.0)
#define RWAT ((K_NAVGDR * K_BOLTZK / K_AMW) * 1000.0)
#define CPD (3.5 * RGAS)
#define CVD (2.5 * RGAS)
#define RCPD (1.0 / CPD)
#define ROVCP (RGAS * RCPD)
#define RDRCV (RGAS / CVD)
#define CPOVR (CPD / RGAS)
#define WLHV 2500610.0
#define TZERO 273.15
#define P00 1.0e5
#define LRATE 0.00649
#define MATHPI 3.1415926535897932384626433832795029
#define DEGRAD (MATHPI / 180.0)
#define RADDEG (180.0 / MATHPI)
#define EARTHRAD 6.371229e6
#define EOMEG 7.2921159e-5
#define EOMEG2 (2.0 * EOMEG)
#define EP1 (K_AMD / K_AMW - 1.0)
#define EP2 (K_AMW / K_AMD)
#define REGRAV (1.0 / K_EGRAV)

/* module parameters of nompi/mod_moloch.F90:97-107 */
#define MINDEN 1.0e-30
#define XDAMP 0.0625
#define NUPAITQ 0.05

void ora_get_constants(ora_constants *c) {
  (void)k_rgasmol;
  (void)k_rgas;
  (void)k_rwat;
  c->egrav = K_EGRAV;
  c->rgas = RGAS;
  c->rwat = RWAT;
  c->cpd = CPD;
  c->cvd = CVD;
  c->rdrcv = RDRCV;
  c->cpovr = CPOVR;
  c->rovcp = ROVCP;
  c->p00 = P00;
  c->ep1 = EP1;
  c->ep2 = EP2;
  c->wlhv = WLHV;
  c->tzero = TZERO;
  c->lrate = LRATE;
  c->eomeg2 = EOMEG2;
  c->degrad = DEGRAD;
  c->mathpi = MATHPI;
  c->regrav = REGRAV;
  c->earthrad = EARTHRAD;
  c->raddeg = RADDEG;
}

/* ------------------------------------------------------------------ layout */
#define PJ (S->nj + 4)
#define PLANE ((long)(S->nj + 4) * (long)(S->ni + 4))
#define I2(j, i) ((long)((i) + 1) * PJ + ((j) + 1))
#define I3(j, i, k) ((long)((k)-1) * PLANE + I2(j, i))
#define F3(a, j, i, k) (a)[I3(j, i, k)]
#define F2(a, j, i) (a)[I2(j, i)]
#define QX(n) (S->qx + (long)((n)-1) * (long)S->nk * PLANE) /* n 1-based */

/* ------------------------------------------------------------- saturation
 * pfesat: nompi/mod_moloch.F90:119-164 (Flatau et al. 1992 polynomials).
 */
double ora_pfesat(double t, double p) {
  const double a0 = 0.611213476e+03, a1 = 0.444007856e+02,
               a2 = 0.143064234e+01, a3 = 0.264461437e-01,
               a4 = 0.305903558e-03, a5 = 0.196237241e-05,
               a6 = 0.892344772e-08, a7 = -0.373208410e-10,
               a8 = 0.209339997e-13;
  const double c0 = 0.611123516e+03, c1 = 0.503109514e+02,
               c2 = 0.188369801e+01, c3 = 0.420547422e-01,
               c4 = 0.614396778e-03, c5 = 0.602780717e-05,
               c6 = 0.387940929e-07, c7 = 0.149436277e-09,
               c8 = 0.262655803e-12;
  double tl = t - TZERO, td, esat;
  if (tl > 100.0) tl = 100.0;
  if (tl < -75.0) tl = -75.0;
  td = tl;
  if (td >= 0.0) {
    esat = a0 + td * (a1 + td * (a2 + td * (a3 + td * (a4 + td * (a5 + td * (a6 + td * (a7 + td * a8)))))));
  } else {
    esat = c0 + td * (c1 + td * (c2 + td * (c3 + td * (c4 + td * (c5 + td * (c6 + td * (c7 + td * c8)))))));
  }
  {
    double cap = 0.15 * p;
    return esat < cap ? esat : cap;
  }
}

/* pfwsat: nompi/mod_moloch.F90:166-181 */
double ora_pfwsat(double t, double p) {
  double es = ora_pfesat(t, p);
  return EP2 * (es / (p - es));
}

/* gzita: mod_zita.F90:110-120 (ratio**2 and ratio**3 as repeated products) */
double ora_gzita(double zita, double ztop, double a0) {
  double ratio = zita / ztop;
  return ((0.0 - 1.0 * a0) * ratio - (3.0 - 2.0 * a0) * (ratio * ratio) +
          (2.0 - 1.0 * a0) * (ratio * ratio * ratio)) +
         1.0;
}

/* ---------------------------------------------------------------- init
 * allocate_moloch xknu (:206-210) and init_moloch (:219-268).
 */
void ora_init_moloch(ora_state *S) {
  int nk = S->nk, k, i, j;
  for (k = 1; k <= nk; k++) {
    S->xknu[k - 1] = XDAMP + (1.0 - XDAMP) *
                                 sin(0.5 * MATHPI * (1.0 - (double)(k - 1) / (double)(nk - 1)));
  }
  for (j = 1; j <= S->nj; j++)
    for (i = 1; i <= S->ni; i++) {
      F2(S->coru, j, i) = EOMEG2 * sin(F2(S->ulat, j, i) * DEGRAD);
      F2(S->corv, j, i) = EOMEG2 * sin(F2(S->vlat, j, i) * DEGRAD);
      F2(S->mx2, j, i) = F2(S->mx, j, i) * F2(S->mx, j, i);
    }
  /* rmu over msfu's bounds (jde1ga:jde2ga, ide1:ide2), rmv over msfv's */
  for (j = 0; j <= S->nj + 1; j++)
    for (i = 1; i <= S->ni; i++) F2(S->rmu, j, i) = 1.0 / F2(S->mu, j, i);
  for (j = 1; j <= S->nj; j++)
    for (i = 0; i <= S->ni + 1; i++) F2(S->rmv, j, i) = 1.0 / F2(S->mv, j, i);
  for (k = 1; k <= nk + 1; k++) S->gzitak[k - 1] = ora_gzita(S->zita[k - 1], S->ztop, S->a0);
  for (k = 1; k <= nk; k++) S->gzitakh[k - 1] = ora_gzita(S->zitah[k - 1], S->ztop, S->a0);
  for (j = 1; j <= S->nj; j++)
    for (i = 1; i <= S->ni; i++) F3(S->w, j, i, 1) = 0.0;
}

/* -------------------------------------------------------------- halo fills
 * The nompi tree replaces every MPI exchange by these periodic copies.
 */
/* 1-wide faces, no corners: e.g. zdiv2 (:859-862), tetav (:708-711),
 * pai/deltaw (:716-723). */
static void halo_faces1(const ora_state *S, double *f, int nlev) {
  int nj = S->nj, ni = S->ni, j, i, k;
  for (k = 1; k <= nlev; k++) {
    for (i = 1; i <= ni; i++) {
      F3(f, 0, i, k) = F3(f, nj, i, k);
      F3(f, nj + 1, i, k) = F3(f, 1, i, k);
    }
    for (j = 1; j <= nj; j++) {
      F3(f, j, 0, k) = F3(f, j, ni, k);
      F3(f, j, ni + 1, k) = F3(f, j, 1, k);
    }
  }
}

/* 1-wide with the four corners: u and v at the start of a substep
 * (:514-529). */
static void halo_corners1(const ora_state *S, double *f) {
  int nj = S->nj, ni = S->ni, k;
  halo_faces1(S, f, S->nk);
  for (k = 1; k <= S->nk; k++) {
    F3(f, nj + 1, ni + 1, k) = F3(f, 1, 1, k);
    F3(f, 0, ni + 1, k) = F3(f, nj, 1, k);
    F3(f, 0, 0, k) = F3(f, nj, ni, k);
    F3(f, nj + 1, 0, k) = F3(f, 1, ni, k);
  }
}

/* 2-wide in j only, rows i=1..ni: u before uvstagtox (:1395-1398), ux
 * before xtouvstag (:1365-1368), p0 in wafone (:1133-1136). */
static void halo_j2(const ora_state *S, double *f) {
  int nj = S->nj, ni = S->ni, i, k;
  for (k = 1; k <= S->nk; k++)
    for (i = 1; i <= ni; i++) {
      F3(f, 0, i, k) = F3(f, nj, i, k);
      F3(f, nj + 1, i, k) = F3(f, 1, i, k);
      F3(f, -1, i, k) = F3(f, nj - 1, i, k);
      F3(f, nj + 2, i, k) = F3(f, 2, i, k);
    }
}

/* 2-wide in i only, columns j=1..nj: v before uvstagtox (:1399-1402), vx
 * before xtouvstag (:1369-1372), wz in wafone (:1075-1078). */
static void halo_i2(const ora_state *S, double *f) {
  int nj = S->nj, ni = S->ni, j, k;
  for (k = 1; k <= S->nk; k++)
    for (j = 1; j <= nj; j++) {
      F3(f, j, 0, k) = F3(f, j, ni, k);
      F3(f, j, ni + 1, k) = F3(f, j, 1, k);
      F3(f, j, -1, k) = F3(f, j, ni - 1, k);
      F3(f, j, ni + 2, k) = F3(f, j, 2, k);
    }
}

/* ------------------------------------------------------- reset_tendencies
 * :1313-1325 */
void ora_reset_tendencies(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  ORA_PAR_IJ
  for (k = 1; k <= nk + 1; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) F3(S->s, j, i, k) = 0.0;
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 0; i <= ni + 1; i++)
      for (j = 0; j <= nj + 1; j++) F3(S->zdiv2, j, i, k) = 0.0;
  ORA_PAR_IJ
  for (k = 1; k <= nk + 1; k++)
    for (i = 0; i <= ni + 1; i++)
      for (j = 0; j <= nj + 1; j++) F3(S->deltaw, j, i, k) = 0.0;
}

/* ---------------------------------------------------------------- prologue
 * :291-323 */
void ora_prologue(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *qv = QX(1), *qc = QX(2), *qi = QX(3), *qr = QX(4), *qs = QX(5);
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->p[x] = pow(S->pai[x], CPOVR) * P00;
        S->qsat[x] = ora_pfwsat(S->t[x], S->p[x]);
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->tvirt[x] = S->t[x] * (1.0 + EP1 * qv[x] - qc[x] - qi[x] - qr[x] - qs[x]);
      }
  if (S->do_fulleq) {
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) {
          long x = I3(j, i, k);
          S->qwltot[x] = qc[x] + qr[x];
          S->qwitot[x] = qi[x] + qs[x];
        }
  }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->tetav[x] = S->tvirt[x] / S->pai[x];
      }
  if (S->do_filterpai) {
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) F3(S->pf, j, i, k) = F3(S->pai, j, i, k);
  }
  if (S->do_fulleq && S->do_filtertheta) {
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) F3(S->tf, j, i, k) = F3(S->tetav, j, i, k);
  }
}

/* ----------------------------------------------------------------- filters
 * divergence_filter :423-438, filtpai :440-460, filttheta :462-481.
 * The 5-point operator 1/8*sum(4 neighbours) - 1/2*centre, added per level.
 */
static void smooth5(ora_state *S, double *f, int k, double coef) {
  int nj = S->nj, ni = S->ni, j, i;
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++)
      F2(S->p2d, j, i) = 0.125 * (F3(f, j - 1, i, k) + F3(f, j + 1, i, k) +
                                  F3(f, j, i - 1, k) + F3(f, j, i + 1, k)) -
                         0.5 * F3(f, j, i, k);
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) F3(f, j, i, k) = F3(f, j, i, k) + coef * F2(S->p2d, j, i);
}

void ora_divergence_filter(ora_state *S) {
  int k;
  halo_faces1(S, S->zdiv2, S->nk);
  for (k = 1; k <= S->nk; k++) {
    /* mo_anu2 * xknu(k) * p2d */
    int nj = S->nj, ni = S->ni, j, i;
    double *f = S->zdiv2;
    ORA_PAR_J
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F2(S->p2d, j, i) = 0.125 * (F3(f, j - 1, i, k) + F3(f, j + 1, i, k) +
                                    F3(f, j, i - 1, k) + F3(f, j, i + 1, k)) -
                           0.5 * F3(f, j, i, k);
    ORA_PAR_J
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(f, j, i, k) = F3(f, j, i, k) + S->anu2 * S->xknu[k - 1] * F2(S->p2d, j, i);
  }
}

void ora_filtpai(ora_state *S) {
  int k;
  halo_faces1(S, S->pai, S->nk);
  for (k = 1; k <= S->nk; k++) smooth5(S, S->pai, k, NUPAITQ);
}

void ora_filttheta(ora_state *S) {
  int k;
  halo_faces1(S, S->tetav, S->nk);
  for (k = 1; k <= S->nk; k++) smooth5(S, S->tetav, k, NUPAITQ);
}

/* ----------------------------------------------------------------- divdamp
 * :853-895 */
void ora_divdamp(ora_state *S, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double ddamp;
  halo_faces1(S, S->zdiv2, nk);
  ddamp = 0.125 * (S->dx / dts);
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(S->u, j, i, k) = F3(S->u, j, i, k) +
                            S->xknu[k - 1] * ddamp * F2(S->mu, j, i) *
                                (F3(S->zdiv2, j, i, k) - F3(S->zdiv2, j - 1, i, k));
  if (S->rotllr) {
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++)
          F3(S->v, j, i, k) = F3(S->v, j, i, k) +
                              S->xknu[k - 1] * ddamp *
                                  (F3(S->zdiv2, j, i, k) - F3(S->zdiv2, j, i - 1, k));
  } else {
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++)
          F3(S->v, j, i, k) = F3(S->v, j, i, k) +
                              S->xknu[k - 1] * ddamp * F2(S->mv, j, i) *
                                  (F3(S->zdiv2, j, i, k) - F3(S->zdiv2, j, i - 1, k));
  }
  ddamp = XDAMP * 0.015625 / dts;
  for (k = 1; k <= nk; k++) smooth5(S, S->zdiv2, k, ddamp);
}

/* -------------------------------------------------------------- thomas
 * Elimination of :664-686 on one column (see header). */
void ora_thomas_column(int nk, const double *zwexpl, const double *zu,
                       const double *zd, double w1, double wkp1, double *w) {
  double wwkw[1024 + 2];
  double wk[1024 + 2];
  int k;
  wk[nk] = wkp1;       /* w(kzp1), index kz (0-based) */
  wwkw[nk] = 0.0;      /* wwkw(j,i,kzp1) = 0 (:508-510) */
  for (k = nk; k >= 2; k--) {
    double zrapp = 1.0 / (1.0 + zd[k - 1] + zu[k - 1] - zd[k - 1] * wwkw[k]);
    wk[k - 1] = zrapp * (zwexpl[k - 1] + zd[k - 1] * wk[k]);
    wwkw[k - 1] = zrapp * zu[k - 1];
  }
  wk[0] = w1;
  for (k = 2; k <= nk; k++) wk[k - 1] = wk[k - 1] + wwkw[k - 1] * wk[k - 2];
  for (k = 2; k <= nk; k++) w[k - 1] = wk[k - 1];
}

/* ------------------------------------------------------------------- sound
 * :483-851.  One substep is the sequence of the phases below (each exposed
 * for the unit tests, which call them on constructed states). */

/* surface boundary condition of w (:533-545), s(kz+1) (:547-549) and the
 * generalised vertical velocity, Equation 10 (:553-570) */
void ora_sound_eq10(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *u = S->u, *v = S->v, *w = S->w, *s = S->s;
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      double zuh = F3(u, j, i, nk) * F2(S->hx, j, i) + F3(u, j + 1, i, nk) * F2(S->hx, j + 1, i);
      double zvh = F3(v, j, i, nk) * F2(S->hy, j, i) + F3(v, j, i + 1, nk) * F2(S->hy, j, i + 1);
      F3(w, j, i, nk + 1) = 0.5 * (zuh + zvh);
    }
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) F3(s, j, i, nk + 1) = -F3(w, j, i, nk + 1);
  ORA_PAR_IJ
  for (k = 2; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double zuh = (F3(u, j, i, k) + F3(u, j, i, k - 1)) * F2(S->hx, j, i) +
                     (F3(u, j + 1, i, k) + F3(u, j + 1, i, k - 1)) * F2(S->hx, j + 1, i);
        double zvh = (F3(v, j, i, k) + F3(v, j, i, k - 1)) * F2(S->hy, j, i) +
                     (F3(v, j, i + 1, k) + F3(v, j, i + 1, k - 1)) * F2(S->hy, j, i + 1);
        F3(s, j, i, k) = -0.25 * (zuh + zvh) * S->gzitak[k - 1];
      }
}

/* horizontal divergence, Equation 16 (:575-624) */
void ora_sound_hdiv(ora_state *S, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double dtrdx = dts / S->dx, dtrdy = dts / S->dx;
  double *u = S->u, *v = S->v, *fmz = S->fmz, *zdiv2 = S->zdiv2;
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double f0 = F3(fmz, j, i, k);
        double zrfmzum = 2.0 / (f0 + F3(fmz, j - 1, i, k));
        double zrfmzvm = 2.0 / (f0 + F3(fmz, j, i - 1, k));
        double zrfmzup = 2.0 / (f0 + F3(fmz, j + 1, i, k));
        double zrfmzvp = 2.0 / (f0 + F3(fmz, j, i + 1, k));
        double zum, zup, zvm, zvp;
        if (S->rotllr) {
          zum = dtrdx * F3(u, j, i, k) * zrfmzum;
          zup = dtrdx * F3(u, j + 1, i, k) * zrfmzup;
          zvm = dtrdy * F3(v, j, i, k) * zrfmzvm * F2(S->rmv, j, i);
          zvp = dtrdy * F3(v, j, i + 1, k) * zrfmzvp * F2(S->rmv, j, i + 1);
          F3(zdiv2, j, i, k) = f0 * F2(S->mx, j, i) * ((zup - zum) + (zvp - zvm));
        } else {
          zum = dtrdx * F3(u, j, i, k) * F2(S->rmu, j, i) * zrfmzum;
          zup = dtrdx * F3(u, j + 1, i, k) * F2(S->rmu, j + 1, i) * zrfmzup;
          zvm = dtrdy * F3(v, j, i, k) * F2(S->rmv, j, i) * zrfmzvm;
          zvp = dtrdy * F3(v, j, i + 1, k) * F2(S->rmv, j, i + 1) * zrfmzvp;
          F3(zdiv2, j, i, k) = F2(S->mx2, j, i) * f0 * ((zup - zum) + (zvp - zvm));
        }
      }
}

/* vertical part of the divergence (:627-630) */
void ora_sound_vdiv(ora_state *S, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double dtrdz = dts / S->dzita;
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(S->zdiv2, j, i, k) = F3(S->zdiv2, j, i, k) +
                                F3(S->fmz, j, i, k) * dtrdz * (F3(S->s, j, i, k) - F3(S->s, j, i, k + 1));
}

/* The coefficients of the implicit w equation of column (j,i) at substep
 * ns (:638-676): zwexpl, zu, zd at index k-1 for k = 2..kz; zrom1w (may be
 * NULL) receives the interface coefficient after the gravity-wave and moist
 * terms.  Also sets deltaw(k) = -w(k) as the reference does there. */
void ora_w_coefficients(ora_state *S, int j, int i, int ns, double dts, double *zwexpl, double *zu,
                        double *zd, double *zrom1w_out) {
  int nk = S->nk, k;
  double dtrdz = dts / S->dzita;
  double zcs2 = dtrdz * dtrdz * RDRCV;
  double *w = S->w, *pai = S->pai, *tetav = S->tetav, *zdiv2 = S->zdiv2, *fmz = S->fmz;
  double *fmzf = S->fmzf, *qv = QX(1);
  for (k = nk; k >= 2; k--) {
    double zrom1w, zqs, zdth, zwe;
    F3(S->deltaw, j, i, k) = -F3(w, j, i, k);
    zrom1w = 0.5 * CPD * F3(fmzf, j, i, k) * (F3(tetav, j, i, k - 1) + F3(tetav, j, i, k));
    zrom1w = zrom1w - CPD * F3(w, j, i, k) * F3(fmzf, j, i, k) * F3(fmzf, j, i, k) * (double)ns * dtrdz *
                          (F3(tetav, j, i, k - 1) - F3(tetav, j, i, k));
    if (F3(qv, j, i, k) > 0.96 * F3(S->qsat, j, i, k) && F3(w, j, i, k) > 0.1) {
      zqs = 0.5 * (F3(S->qsat, j, i, k) + F3(S->qsat, j, i, k - 1));
      zdth = K_EGRAV * F3(w, j, i, k) * (double)(ns - 1) * dts * WLHV * WLHV * zqs /
             (CPD * F3(pai, j, i, k - 1) * RWAT * F3(S->t, j, i, k - 1) * F3(S->t, j, i, k - 1));
      zrom1w = zrom1w + zdth * F3(fmzf, j, i, k);
    }
    if (zrom1w_out) zrom1w_out[k - 1] = zrom1w;
    zwe = F3(w, j, i, k) - zrom1w * dtrdz * (F3(pai, j, i, k - 1) - F3(pai, j, i, k)) - K_EGRAV * dts;
    zwe = zwe + RDRCV * zrom1w * dtrdz *
                    (F3(pai, j, i, k - 1) * F3(zdiv2, j, i, k - 1) - F3(pai, j, i, k) * F3(zdiv2, j, i, k));
    zwexpl[k - 1] = zwe;
    zu[k - 1] = zcs2 * F3(fmz, j, i, k - 1) * zrom1w * F3(pai, j, i, k - 1) + S->ffilt[k - 1];
    zd[k - 1] = zcs2 * F3(fmz, j, i, k) * zrom1w * F3(pai, j, i, k) + S->ffilt[k - 1];
  }
}

/* implicit w, Equation 19 (:634-686), one column at a time, and deltaw */
void ora_sound_wsolve(ora_state *S, int ns, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double zwexpl[1026], zu[1026], zd[1026], wcol[1026];
  double *w = S->w;
  ORA_PAR_COLUMNS
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      ora_w_coefficients(S, j, i, ns, dts, zwexpl, zu, zd, NULL);
      ora_thomas_column(nk, zwexpl, zu, zd, F3(w, j, i, 1), F3(w, j, i, nk + 1), wcol);
      for (k = 2; k <= nk; k++) {
        F3(w, j, i, k) = wcol[k - 1];
        F3(S->deltaw, j, i, k) = F3(S->deltaw, j, i, k) + F3(w, j, i, k);
      }
    }
}

/* The fulleq moist corrections of one point (:697-706): zdiv2 scaled by the
 * moist factor, then tetav from the scaled zdiv2. */
void ora_fulleq_point(double qv, double qc, double qwltot, double qwitot, double *zdiv2, double *tetav) {
  *zdiv2 = *zdiv2 * (1.0 + 0.86 * qv + 3.2 * qc) / (1.0 + 0.96 * qv + 4.8 * qc);
  *tetav = *tetav * (1.0 + RDRCV * *zdiv2 * (0.25 * qv + 4.2 * qwltot + 2.1 * qwitot));
}

/* w part of the divergence, the fulleq corrections, the divergence filter,
 * the halos and the new Exner function (:690-728) */
void ora_sound_exner(ora_state *S, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double dtrdz = dts / S->dzita;
  double *w = S->w, *pai = S->pai, *zdiv2 = S->zdiv2;
  double *qv = QX(1), *qc = QX(2);
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(zdiv2, j, i, k) = F3(zdiv2, j, i, k) + dtrdz * F3(S->fmz, j, i, k) * (F3(w, j, i, k) - F3(w, j, i, k + 1));

  if (S->do_fulleq) { /* :695-712 */
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) {
          long x = I3(j, i, k);
          ora_fulleq_point(qv[x], qc[x], S->qwltot[x], S->qwitot[x], &zdiv2[x], &S->tetav[x]);
        }
    halo_faces1(S, S->tetav, nk);
  }

  if (S->divfilter) ora_divergence_filter(S); /* :714 */

  if (!S->halo_mpi) { /* nompi: halos taken before the Exner update (:716-723) */
    halo_faces1(S, pai, nk);
    halo_faces1(S, S->deltaw, nk + 1);
  }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++) /* :726-728 */
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) F3(pai, j, i, k) = F3(pai, j, i, k) * (1.0 - RDRCV * F3(zdiv2, j, i, k));
  if (S->halo_mpi) { /* MPI tree order, mod_moloch.F90:697-703 */
    halo_faces1(S, pai, nk);
    halo_faces1(S, S->deltaw, nk + 1);
  }
}

/* snapshots including halos (:730-735) and the horizontal momentum,
 * Equations 17 and 18 (:737-838) */
void ora_sound_momentum(ora_state *S, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double dtrdx = dts / S->dx, dtrdy = dts / S->dx;
  double *u = S->u, *v = S->v, *pai = S->pai, *tetav = S->tetav, *deltaw = S->deltaw;
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 0; i <= ni + 1; i++)
      for (j = 0; j <= nj + 1; j++) {
        F3(S->ud, j, i, k) = F3(u, j, i, k);
        F3(S->vd, j, i, k) = F3(v, j, i, k);
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double zcx = dtrdx * F2(S->mu, j, i);
        double zfz = K_EGRAV * dts + 0.25 * (F3(deltaw, j - 1, i, k) + F3(deltaw, j - 1, i, k + 1) +
                                             F3(deltaw, j, i, k) + F3(deltaw, j, i, k + 1));
        double zrom1u = 0.5 * CPD * (F3(tetav, j - 1, i, k) + F3(tetav, j, i, k));
        double zcor1u = F2(S->coru, j, i) * dts * 0.25 *
                        (F3(S->vd, j, i, k) + F3(S->vd, j - 1, i, k) + F3(S->vd, j - 1, i + 1, k) +
                         F3(S->vd, j, i + 1, k));
        F3(u, j, i, k) = F3(u, j, i, k) + zcor1u - zfz * F2(S->hx, j, i) * S->gzitakh[k - 1] -
                         zcx * zrom1u * (F3(pai, j, i, k) - F3(pai, j - 1, i, k));
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double zcy = S->rotllr ? dtrdy : dtrdy * F2(S->mv, j, i);
        double zfz = K_EGRAV * dts + 0.25 * (F3(deltaw, j, i - 1, k) + F3(deltaw, j, i - 1, k + 1) +
                                             F3(deltaw, j, i, k) + F3(deltaw, j, i, k + 1));
        double zrom1v = 0.5 * CPD * (F3(tetav, j, i - 1, k) + F3(tetav, j, i, k));
        double zcor1v = F2(S->corv, j, i) * dts * 0.25 *
                        (F3(S->ud, j, i, k) + F3(S->ud, j, i - 1, k) + F3(S->ud, j + 1, i, k) +
                         F3(S->ud, j + 1, i - 1, k));
        F3(v, j, i, k) = F3(v, j, i, k) - zcor1v - zfz * F2(S->hy, j, i) * S->gzitakh[k - 1] -
                         zcy * zrom1v * (F3(pai, j, i, k) - F3(pai, j, i - 1, k));
      }
}

/* One acoustic substep ns of length dts (:512-838). */
void ora_substep(ora_state *S, int ns, double dts) {
  halo_corners1(S, S->u); /* :514-521 */
  halo_corners1(S, S->v); /* :522-529 */
  ora_sound_eq10(S);
  ora_sound_hdiv(S, dts);
  if (S->do_divdamp) ora_divdamp(S, dts); /* :626 */
  ora_sound_vdiv(S, dts);
  ora_sound_wsolve(S, ns, dts);
  ora_sound_exner(S, dts);
  ora_sound_momentum(S, dts);
}

void ora_sound(ora_state *S, double dts) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k, ns;
  double *s = S->s, *w = S->w;

  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) F3(S->wwkw, j, i, nk + 1) = 0.0;

  if (!S->do_fulleq) halo_faces1(S, S->tetav, nk);

  for (ns = 1; ns <= S->nsound; ns++) ora_substep(S, ns, dts);

  /* complete Equation 10 (:843-849) */
  ORA_PAR_IJ
  for (k = 2; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) F3(s, j, i, k) = (F3(w, j, i, k) + F3(s, j, i, k)) * F3(S->fmzf, j, i, k);
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      F3(s, j, i, 1) = 0.0;
      F3(s, j, i, nk + 1) = 0.0;
    }
}

/* ------------------------------------------------------------ WAF limiter
 * superbee in the form used at :990-993 */
static inline double waf_b(double r) {
  double m = fmin(2.0 * r, 1.0);
  m = fmax(r, m);
  m = fmin(2.0, m);
  return fmax(0.0, m);
}
static inline double sgn_floor(double z) { return copysign(fmax(fabs(z), MINDEN), z); }

/* ------------------------------------------------------------------ wafone
 * :929-1311 (pfac/pmin optional arguments are never passed by advection) */
void ora_wafone(ora_state *S, double *pp, double dta) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k, pass;
  double dtrdx = dta / S->dx, dtrdy = dta / S->dx, dtrdz = dta / S->dzita;
  double *s = S->s, *fmz = S->fmz, *fmzf = S->fmzf, *wz = S->wz, *p0 = S->p0;
  double *wfw = S->wfw, *zpby = S->zpby, *zpbw = S->zpbw, *u = S->u, *v = S->v;
  const int imin = -1, imax = ni + 2, jmin = -1, jmax = nj + 2; /* :264-267 */
  if (S->do_vadvtwice) dtrdz = 0.5 * dtrdz;

  /* vertical advection (:951-1073): pass 0 reads pp, pass 1 reads wz */
  for (pass = 0; pass < (S->do_vadvtwice ? 2 : 1); pass++) {
    double *q = pass == 0 ? pp : wz;
    ORA_PAR_J
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        F3(wfw, j, i, 1) = 0.0;
        F3(wfw, j, i, nk + 1) = 0.0;
      }
    ORA_PAR_IJ
    for (k = 1; k <= nk - 1; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) {
          double zamu = F3(s, j, i, k + 1) * dtrdz, is, zzden, r, b, zphi;
          int k1, k1p1;
          if (zamu >= 0.0) {
            is = 1.0;
            k1 = k + 1;
            k1p1 = k1 + 1;
            if (k1p1 > nk) k1p1 = nk;
          } else {
            is = -1.0;
            k1 = k - 1;
            k1p1 = k;
            if (k1 < 1) k1 = 1;
          }
          zzden = sgn_floor(F3(q, j, i, k) - F3(q, j, i, k + 1));
          r = (F3(q, j, i, k1) - F3(q, j, i, k1p1)) / zzden;
          b = waf_b(r);
          zphi = is + zamu * b - is * b;
          F3(wfw, j, i, k + 1) = 0.5 * F3(s, j, i, k + 1) *
                                 ((1.0 + zphi) * F3(q, j, i, k + 1) + (1.0 - zphi) * F3(q, j, i, k));
        }
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) {
          double zrfmu = dtrdz * F3(fmz, j, i, k) / F3(fmzf, j, i, k);
          double zrfmd = dtrdz * F3(fmz, j, i, k) / F3(fmzf, j, i, k + 1);
          double zdv = (F3(s, j, i, k) * zrfmu - F3(s, j, i, k + 1) * zrfmd) * F3(q, j, i, k);
          F3(wz, j, i, k) = F3(q, j, i, k) - F3(wfw, j, i, k) * zrfmu +
                            F3(wfw, j, i, k + 1) * zrfmd + zdv;
        }
  }

  halo_i2(S, wz); /* :1075-1078 */

  /* meridional advection (:1082-1131 / :1190-1240) */
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni + 1; i++)
      for (j = 1; j <= nj; j++) {
        double vv = F3(v, j, i, k), zamu, is, zzden, r, b, zphi;
        int ih, ihm1;
        zamu = S->rotllr ? vv * dtrdy : vv * F2(S->rmv, j, i) * dtrdy;
        if (zamu > 0.0) {
          is = 1.0;
          ih = i - 1;
        } else {
          is = -1.0;
          ih = (i + 1 < imax) ? i + 1 : imax;
        }
        ihm1 = (ih - 1 > imin) ? ih - 1 : imin;
        zzden = sgn_floor(F3(wz, j, i, k) - F3(wz, j, i - 1, k));
        r = (F3(wz, j, ih, k) - F3(wz, j, ihm1, k)) / zzden;
        b = waf_b(r);
        zphi = is + zamu * b - is * b;
        if (S->rotllr)
          F3(zpby, j, i, k) = 0.5 * vv * ((1.0 + zphi) * F3(wz, j, i - 1, k) + (1.0 - zphi) * F3(wz, j, i, k));
        else
          F3(zpby, j, i, k) = 0.5 * vv * F2(S->rmv, j, i) *
                              ((1.0 + zphi) * F3(wz, j, i - 1, k) + (1.0 - zphi) * F3(wz, j, i, k));
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double f0 = F3(fmz, j, i, k);
        if (S->rotllr) {
          double zhxvtn = dtrdy * F2(S->rmv, j, i + 1) * F2(S->mx, j, i);
          double zhxvts = dtrdy * F2(S->rmv, j, i) * F2(S->mx, j, i);
          double zrfmn = zhxvtn * 2.0 * f0 / (f0 + F3(fmz, j, i + 1, k));
          double zrfms = zhxvts * 2.0 * f0 / (f0 + F3(fmz, j, i - 1, k));
          double zdv = (F3(v, j, i + 1, k) * zrfmn - F3(v, j, i, k) * zrfms) * F3(pp, j, i, k);
          F3(p0, j, i, k) = F3(wz, j, i, k) + F3(zpby, j, i, k) * zrfms -
                            F3(zpby, j, i + 1, k) * zrfmn + zdv;
        } else {
          double zrfmn = dtrdy * 2.0 * f0 / (f0 + F3(fmz, j, i + 1, k));
          double zrfms = dtrdy * 2.0 * f0 / (f0 + F3(fmz, j, i - 1, k));
          double zdv = (F3(v, j, i + 1, k) * F2(S->rmv, j, i + 1) * zrfmn -
                        F3(v, j, i, k) * F2(S->rmv, j, i) * zrfms) *
                       F3(pp, j, i, k);
          F3(p0, j, i, k) = F3(wz, j, i, k) + F2(S->mx2, j, i) * (F3(zpby, j, i, k) * zrfms -
                                                                 F3(zpby, j, i + 1, k) * zrfmn + zdv);
        }
      }

  halo_j2(S, p0); /* :1133-1136 */

  /* zonal advection (:1138-1187 / :1247-1296) */
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj + 1; j++) {
        double uu = F3(u, j, i, k), zamu, is, zzden, r, b, zphi;
        int jh, jhm1;
        zamu = S->rotllr ? uu * F2(S->mu, j, i) * dtrdx : uu * F2(S->rmu, j, i) * dtrdx;
        if (zamu > 0.0) {
          is = 1.0;
          jh = j - 1;
        } else {
          is = -1.0;
          jh = (j + 1 < jmax) ? j + 1 : jmax;
        }
        jhm1 = (jh - 1 > jmin) ? jh - 1 : jmin;
        zzden = sgn_floor(F3(p0, j, i, k) - F3(p0, j - 1, i, k));
        r = (F3(p0, jh, i, k) - F3(p0, jhm1, i, k)) / zzden;
        b = waf_b(r);
        zphi = is + zamu * b - is * b;
        if (S->rotllr)
          F3(zpbw, j, i, k) = 0.5 * uu * ((1.0 + zphi) * F3(p0, j - 1, i, k) + (1.0 - zphi) * F3(p0, j, i, k));
        else
          F3(zpbw, j, i, k) = 0.5 * uu * F2(S->rmu, j, i) *
                              ((1.0 + zphi) * F3(p0, j - 1, i, k) + (1.0 - zphi) * F3(p0, j, i, k));
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double f0 = F3(fmz, j, i, k);
        if (S->rotllr) {
          double zcostx = dtrdx * F2(S->mu, j, i);
          double zrfme = zcostx * 2.0 * f0 / (f0 + F3(fmz, j + 1, i, k));
          double zrfmw = zcostx * 2.0 * f0 / (f0 + F3(fmz, j - 1, i, k));
          double zdv = (F3(u, j + 1, i, k) * zrfme - F3(u, j, i, k) * zrfmw) * F3(pp, j, i, k);
          F3(pp, j, i, k) = F3(p0, j, i, k) + F3(zpbw, j, i, k) * zrfmw -
                            F3(zpbw, j + 1, i, k) * zrfme + zdv;
        } else {
          double zrfme = dtrdx * 2.0 * f0 / (f0 + F3(fmz, j + 1, i, k));
          double zrfmw = dtrdx * 2.0 * f0 / (f0 + F3(fmz, j - 1, i, k));
          double zdv = (F3(u, j + 1, i, k) * F2(S->rmu, j + 1, i) * zrfme -
                        F3(u, j, i, k) * F2(S->rmu, j, i) * zrfmw) *
                       F3(pp, j, i, k);
          F3(pp, j, i, k) = F3(p0, j, i, k) + F2(S->mx2, j, i) * (F3(zpbw, j, i, k) * zrfmw -
                                                                 F3(zpbw, j + 1, i, k) * zrfme + zdv);
        }
      }
}

/* ------------------------------------------------------------ interpolators */
/* wstagtox :1327-1341 */
void ora_wstagtox(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *w = S->w, *wx = S->wx;
  ORA_PAR_IJ
  for (k = 2; k <= nk - 1; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(wx, j, i, k) = 0.5625 * (F3(w, j, i, k + 1) + F3(w, j, i, k)) -
                          0.0625 * (F3(w, j, i, k + 2) + F3(w, j, i, k - 1));
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      F3(wx, j, i, 1) = 0.5 * (F3(w, j, i, 2) + F3(w, j, i, 1));
      F3(wx, j, i, nk) = 0.5 * (F3(w, j, i, nk + 1) + F3(w, j, i, nk));
    }
}

/* xtowstag :1343-1357 */
void ora_xtowstag(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *w = S->w, *wx = S->wx;
  ORA_PAR_IJ
  for (k = 3; k <= nk - 1; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(w, j, i, k) = 0.5625 * (F3(wx, j, i, k) + F3(wx, j, i, k - 1)) -
                         0.0625 * (F3(wx, j, i, k + 1) + F3(wx, j, i, k - 2));
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      F3(w, j, i, 2) = 0.5 * (F3(wx, j, i, 2) + F3(wx, j, i, 1));
      F3(w, j, i, nk) = 0.5 * (F3(wx, j, i, nk) + F3(wx, j, i, nk - 1));
    }
}

/* xtouvstag :1359-1387 */
void ora_xtouvstag(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *ux = S->ux, *vx = S->vx, *u = S->u, *v = S->v;
  halo_j2(S, ux);
  halo_i2(S, vx);
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(u, j, i, k) = 0.5625 * (F3(ux, j, i, k) + F3(ux, j - 1, i, k)) -
                         0.0625 * (F3(ux, j + 1, i, k) + F3(ux, j - 2, i, k));
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(v, j, i, k) = 0.5625 * (F3(vx, j, i, k) + F3(vx, j, i - 1, k)) -
                         0.0625 * (F3(vx, j, i + 1, k) + F3(vx, j, i - 2, k));
}

/* uvstagtox :1389-1417 */
void ora_uvstagtox(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *ux = S->ux, *vx = S->vx, *u = S->u, *v = S->v;
  halo_j2(S, u);
  halo_i2(S, v);
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(ux, j, i, k) = 0.5625 * (F3(u, j + 1, i, k) + F3(u, j, i, k)) -
                          0.0625 * (F3(u, j + 2, i, k) + F3(u, j - 1, i, k));
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++)
        F3(vx, j, i, k) = 0.5625 * (F3(v, j, i + 1, k) + F3(v, j, i, k)) -
                          0.0625 * (F3(v, j, i + 2, k) + F3(v, j, i - 1, k));
}

/* --------------------------------------------------------------- advection
 * :897-927 */
void ora_advection(ora_state *S, double dta) {
  int n;
  ora_uvstagtox(S);
  ora_wstagtox(S);
  ora_wafone(S, S->tetav, dta);
  ora_wafone(S, S->pai, dta);
  ora_wafone(S, S->ux, dta);
  ora_wafone(S, S->vx, dta);
  ora_wafone(S, S->wx, dta);
  ora_wafone(S, QX(1), dta);
  for (n = S->iqfrst; n <= S->nqx; n++) ora_wafone(S, QX(n), dta);
  ora_xtouvstag(S);
  ora_xtowstag(S);
}

/* ---------------------------------------------------------------- epilogue
 * :355-391 */
void ora_epilogue(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double *qv = QX(1), *qc = QX(2), *qi = QX(3), *qr = QX(4), *qs = QX(5);
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->tvirt[x] = S->tetav[x] * S->pai[x];
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->t[x] = S->tvirt[x] / (1.0 + EP1 * qv[x] - qc[x] - qi[x] - qr[x] - qs[x]);
      }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->p[x] = pow(S->pai[x], CPOVR) * P00;
        S->rho[x] = S->p[x] / (RGAS * S->t[x]);
      }
  ORA_PAR_J
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      double zdgz = F3(S->zeta, j, i, nk) * K_EGRAV;
      double lrt = (F3(S->tvirt, j, i, nk - 1) - F3(S->tvirt, j, i, nk)) /
                   (F3(S->zeta, j, i, nk - 1) - F3(S->zeta, j, i, nk));
      double tv;
      lrt = 0.65 * lrt - 0.35 * LRATE;
      tv = F3(S->tvirt, j, i, nk) - 0.5 * F3(S->zeta, j, i, nk) * lrt;
      F2(S->ps, j, i) = F3(S->p, j, i, nk) * exp(zdgz / (RGAS * tv));
    }
  ORA_PAR_IJ
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        long x = I3(j, i, k);
        S->qsat[x] = ora_pfwsat(S->t[x], S->p[x]);
      }
}

/* ----------------------------------------------------------- diagnostics
 * :395-419 (extrema over the interior; NaN/Inf check of ps) */
void ora_diagnostics(ora_state *S) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k;
  double mxp = -INFINITY, mnp = INFINITY, mxu = -INFINITY, mnu = INFINITY;
  double mxv = -INFINITY, mnv = INFINITY;
  int bad = 0;
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++) {
      double x = F2(S->ps, j, i);
      if (isnan(x)) bad = 1;
      if (x > mxp) mxp = x;
      if (x < mnp) mnp = x;
    }
  for (k = 1; k <= nk; k++)
    for (i = 1; i <= ni; i++)
      for (j = 1; j <= nj; j++) {
        double a = F3(S->u, j, i, k), b = F3(S->v, j, i, k);
        if (a > mxu) mxu = a;
        if (a < mnu) mnu = a;
        if (b > mxv) mxv = b;
        if (b < mnv) mnv = b;
      }
  S->diag[0] = mxp;
  S->diag[1] = mnp;
  S->diag[2] = mxu;
  S->diag[3] = mnu;
  S->diag[4] = mxv;
  S->diag[5] = mnv;
  S->cfl_violation = bad || isnan(mxp) || isnan(mnp) || isinf(mxp) || isinf(mnp);
}

/* ------------------------------------------------------------------ moloch
 * :272-421 */
void ora_moloch(ora_state *S, int istep, int lprint) {
  int nj = S->nj, ni = S->ni, nk = S->nk, j, i, k, nadv;
  double dtstepa = S->dt / (double)S->nadv;
  double dtsound = dtstepa / (double)S->nsound;
  (void)istep;

  ora_reset_tendencies(S);
  ora_prologue(S);

  for (nadv = 1; nadv <= S->nadv; nadv++) {
    ora_sound(S, dtsound);
    ora_advection(S, dtstepa);
  }

  if (S->do_filterpai) { /* :333-341 */
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) F3(S->pai, j, i, k) = F3(S->pai, j, i, k) - F3(S->pf, j, i, k);
    ora_filtpai(S);
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++) F3(S->pai, j, i, k) = F3(S->pai, j, i, k) + F3(S->pf, j, i, k);
  }
  if (S->do_fulleq && S->do_filtertheta) { /* :343-353 */
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++)
          F3(S->tetav, j, i, k) = F3(S->tetav, j, i, k) - F3(S->tf, j, i, k);
    ora_filttheta(S);
    ORA_PAR_IJ
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        for (j = 1; j <= nj; j++)
          F3(S->tetav, j, i, k) = F3(S->tetav, j, i, k) + F3(S->tf, j, i, k);
  }

  ora_epilogue(S);
  if (lprint) ora_diagnostics(S);
}
