/*
 * mdx_oracle.c -- CPU FP64 restatement of cook's force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never linked into the engine (macsimus_amd/), and the
 * engine never falls back to it.
 *
 * It restates, in plain C, the reference's algorithm for cook's force call
 * with the compile-time variant COULOMB=-2, SPLINE=-2, LINKCELL (serial),
 * WORM, PRESSURETENSOR=7, nonpolar LJ:
 *   combining rule + smooth cutoff  sim/lj/sitesite.c:66-112, sim/setss.c:137-306
 *   tail correction                 sim/setss.c:455-501
 *   erfc hyperbolic splines         sim/erfc.c:68-191, sim/splines.c:23-158
 *   exact exclusion erf             sim/erexact.c:25-53
 *   Ewald alpha/kappa auto set      sim/ewald.c:1027-1067
 *   k-vector set, almostint         sim/ewald.c:240-250, 484-534
 *   Ewald pass 1 / pass 2m          sim/ewald.c:562-692, sim/ewpass1.c, sim/ewpass2m.c:189-300
 *   dipole term, self term          sim/ewald.c:865-997, sim/ewaldf.c
 *   linked-cell list                cook/linklist.c:15-29, 37-269, 294-337
 *   half-shell traversal            cook/lc.c:217-460
 *   LJ + Coulomb pair kernel        cook/ljqqnp.c:43-259, sim/intermac.h:47-64
 *   SHAKE + leap-frog + Berendsen   sim/constrd.c:622-1081, sim/verlet.c, sim/virc.c,
 *                                   sim/shakev3.c, sim/thermo.c:141-160, sim/constrd.c:98-107
 *   other site-site forms           exp-6 sim/exp6/sitesite.{c,h}, Buckingham sim/buck,
 *                                   Busing sim/busing (POW = 0), all in the exp-6 table
 *   RDF, Nose-Hoover, frictions     cook/ljqqnp.c:218-225, sim/trvpscale.c, sim/thermo.c:107-135
 *   MTK / Hoover isotropic NPT      sim/trvpscale.c:43-70, sim/thermo.c:58-99, sim/norm.c:909-1016
 * (these later functions state their own citations and how they are pinned)
 * Everything is double precision, serial except the k-space sums, which
 * split the charges over OpenMP threads as cook's PARALLEL=1 does
 * (sim/ewald.c:197-244).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/mdx.h"

#define PI 3.14159265358979323846
#define SQRTPI 1.772453850905516

static double sqr(double x) { return x * x; }
static double pow6(double x) { double y = x * x * x; return y * y; }

/* ------------------------------------------------------------------ LJ */

/* sigvdW/epsvdW of a pair (sim/lj/sitesite.c:66-93), square-root energy
   rule (comb_rule&1, or alpha<=0) with additive (0) or geometric (2)
   diameters.  EvdW < 0 in K. */
static void orc_combrule(double *sig, double *eps, double Ei, double Ri,
                         double Ej, double Rj, int comb_rule)
{
  double r, A, B, Esi = sqrt(-2 * Ei), Esj = sqrt(-2 * Ej);
  if (comb_rule & 2) { r = 4 * Ri * Rj; r = r * r * r; }
  else r = pow6(Ri + Rj);
  B = Esi * Esj * r;
  A = B * r;
  *sig = B == 0 ? 0 : pow(A / B, 1. / 6);
  *eps = A == 0 ? 0 : B * B / (-2 * A);
}

static double orc_U(double sig, double eps, double r)
{ double x = pow6(sig / r); return -eps * (x * x - 2 * x); }
static double orc_F(double sig, double eps, double r)
{ double x = pow6(sig / r); return -12 * eps * (x * x - x) / r; }

int oracle_setss(mdx_lj_t *ss, double Ei, double Ri, double Ej, double Rj,
                 int comb_rule, double C2, double *corr)
{
  const double neps = 1.1102230246251565e-16; /* No.eps = machprec() */
  double sig, eps, C1, oldC1;
  int n;
  orc_combrule(&sig, &eps, Ei, Ri, Ej, Rj, comb_rule);
  memset(ss, 0, sizeof(*ss));
  ss->Sq = sig * sig * 0.7937005259840998;
  ss->E4 = -4 * eps;
  if (ss->Sq != 0) { ss->E24 = -24 * eps / ss->Sq; ss->E48 = -48 * eps / ss->Sq; }
  if (eps == 0 || C2 == 0) {           /* no LJ: C1=C2=A=0 */
    if (corr) *corr = 0;
    return 0;                          /* A4 left 0 (cook leaves it NaN, never read) */
  }
  C1 = 0.775 * C2;
  n = 0;
  do {
    double den = orc_F(sig, eps, C1);
    if (fabs(den) < 1e-50) return MDX_E_ARG;
    oldC1 = C1;
    C1 = 1 + 4 * orc_U(sig, eps, C1) / C1 / den;
    if (n++ > 2000 || C1 < 0) return MDX_E_ARG;
    C1 = C2 / sqrt(C1);
  } while (fabs(1 - C1 / oldC1) > neps * n);
  ss->A = orc_U(sig, eps, C1) / sqr(C2 * C2 - C1 * C1);
  ss->A4 = -orc_F(sig, eps, C1) / (C2 * C2 - C1 * C1) / C1;
  ss->C2q = C2 * C2;
  ss->C1q = C1 * C1;
  if (corr) {
    /* 4 pi int_C1^C2 A(r^2-C2^2)^2 r^2 analytically (with minus sign) +
       4 pi int_C1^inf U r^2 by 2-point Gauss in x=1/r, NGauss4 = 64
       intervals (sim/setss.c:49,470-499) */
    const int NG = 64;
    const double g4 = 0.2113248654051871;
    double c0 = 4 * PI * ss->A *
      (C1 * (pow6(C1) / 7 - 0.4 * sqr(C1) * sqr(C1) * sqr(C2) + sqr(C1) * sqr(C2) * sqr(C2) / 3)
       - 8.0 / 105 * pow(C2, 7));
    double h = 1.0 / NG / C1, c1 = 0, r;
    int k;
    for (k = 0; k < NG; k++) {
      r = 1 / ((k + g4) * h); c1 += sqr(r * r) * orc_U(sig, eps, r);
      r = 1 / ((k + (1 - g4)) * h); c1 += sqr(r * r) * orc_U(sig, eps, r);
    }
    *corr = c0 + c1 * 2 * PI * h;
  }
  return 0;
}

/* --------------------------------------------------------------- exp-6 */

/* the exp-6 macros of sim/exp6/sitesite.h:59-75 evaluated from rr = r^2
   with the constants of initssaux() (sim/exp6/sitesite.c:78-85): U, and
   f*r, the force, as ss_u() and ss_f() return them (sim/setss.c:105-134) */
typedef struct { double tau, eu, ef, ru, rf; } orc_e6;

static double orc_e6U(const orc_e6 *p, double r)
{ double rr = r * r, x = sqrt(rr); return p->eu * exp(p->tau * x) - p->ru / (rr * rr * rr); }
static double orc_e6F(const orc_e6 *p, double r)
{
  double rr = r * r, x = sqrt(rr), z = rr * rr * rr;
  return (p->ef * exp(p->tau * x) / x - p->rf / z / rr) * r;
}

/* the potential-independent smooth cutoff and tail correction of
   setss() (sim/setss.c:244-306, 455-499) on the exp-6 u and f of p */
static int orc_e6_cut(const orc_e6 *q, mdx_exp6_t *ss, double C2, double *corr)
{
  const double neps = 1.1102230246251565e-16;
  orc_e6 p = *q;
  double C1, oldC1;
  int n;
  ss->tau = p.tau; ss->eu = p.eu; ss->ef = p.ef; ss->ru = p.ru;
  C1 = 0.775 * C2;
  n = 0;
  do {
    double den = orc_e6F(&p, C1);
    if (fabs(den) < 1e-50) return MDX_E_ARG;
    oldC1 = C1;
    C1 = 1 + 4 * orc_e6U(&p, C1) / C1 / den;
    if (n++ > 2000 || C1 < 0) return MDX_E_ARG;
    C1 = C2 / sqrt(C1);
  } while (fabs(1 - C1 / oldC1) > neps * n);
  if (C1 > 0.999 * C2 || C1 < 0.5 * C2) return MDX_E_ARG;
  ss->A = orc_e6U(&p, C1) / sqr(C2 * C2 - C1 * C1);
  ss->A4 = -orc_e6F(&p, C1) / (C2 * C2 - C1 * C1) / C1;
  ss->C2q = C2 * C2;
  ss->C1q = C1 * C1;
  if (corr) {
    const int NG = 64;
    const double g4 = 0.2113248654051871;
    double c0 = 4 * PI * ss->A *
      (C1 * (pow6(C1) / 7 - 0.4 * sqr(C1) * sqr(C1) * sqr(C2) + sqr(C1) * sqr(C2) * sqr(C2) / 3)
       - 8.0 / 105 * pow(C2, 7));
    double h = 1.0 / NG / C1, c1 = 0, r;
    int k;
    for (k = 0; k < NG; k++) {
      r = 1 / ((k + g4) * h); c1 += sqr(r * r) * orc_e6U(&p, r);
      r = 1 / ((k + (1 - g4)) * h); c1 += sqr(r * r) * orc_e6U(&p, r);
    }
    *corr = c0 + c1 * 2 * PI * h;
  }
  return 0;
}

/* combrule() of sim/exp6/sitesite.c:60-76 and the potential-independent
   smooth cutoff and tail correction of setss() (sim/setss.c:244-306,
   455-499) on the exp-6 u and f */
int oracle_setss_exp6(mdx_exp6_t *ss, double Ei, double Ri, double ai, double Ej,
                      double Rj, double aj, int comb_rule, double C2, double *corr)
{
  double sig, eps, a;
  orc_e6 p;
  memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (!(comb_rule & 1) || Ei > 0 || Ej > 0) return MDX_E_ARG;
  sig = comb_rule & 2 ? sqrt(Ri * Rj * 4) : Ri + Rj;
  eps = sqrt(Ei * Ej);
  a = sqrt(ai * aj);
  if (eps == 0 || C2 == 0) return 0;
  p.tau = -(a / sig);
  p.eu = 6 * eps / (a - 6) * exp(a);
  p.ef = p.eu * (a / sig);
  p.ru = eps * a * pow6(sig) / (a - 6);
  p.rf = 6 * p.ru;
  return orc_e6_cut(&p, ss, C2, corr);
}

/* Buckingham (sim/buck, POW = 0): u = A_ij exp(-B_ij r) - C_ij/r^6, the
   exp-6 macros with tau = -B_ij, eu = A_ij, ef = A_ij B_ij, ru = C_ij
   (initssaux(), sim/buck/sitesite.c:178-188); the combining rules of
   combrule() (sim/buck/sitesite.c:113-170); setss() as for exp-6 */
int oracle_setss_buck(mdx_exp6_t *ss, double Ai, double Bi, double Ci, double Aj,
                      double Bj, double Cj, int comb_rule, double C2, double *corr)
{
  double A = 0, B = 0, aux = 0, C;
  orc_e6 p;
  memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (((comb_rule & 12) == 0) != ((comb_rule & 3) == 0)) return MDX_E_ARG;
  if (Ai < 0 || Aj < 0 || Bi < 0 || Bj < 0 || Ci < 0 || Cj < 0) return MDX_E_ARG;
  if ((comb_rule & 3) == 2 || Bi * Bj != 0) switch (comb_rule & 3) {
    case 0: aux = 0.5 / pow6(Bi) + 0.5 / pow6(Bj); B = pow(aux, -1. / 6); break;
    case 1: B = sqrt(Bi) * sqrt(Bj); break;
    case 2: B = (Bi + Bj) / 2; break;
    case 3: B = 1 / (0.5 / Bi + 0.5 / Bj); break;
  }
  if (Ai * Aj != 0) switch (comb_rule & 12) {
    case 0: A = sqrt(Ai) / (Bi * Bi * Bi) * (sqrt(Aj) / (Bj * Bj * Bj)) / aux; break;
    case 4: A = sqrt(Ai) * sqrt(Aj); break;
    case 8: {
      double qi = Bi / (Bi + Bj), x = Aj * Bj / (Ai * Bi);
      A = (Ai * pow(x, qi) + Aj * pow(x, qi - 1)) / 2;
      break;
    }
    case 12: {
      double qi = (0.5 / Bi) / (0.5 / Bi + 0.5 / Bj);
      A = pow(Ai * Bi, qi) * pow(Aj * Bj, 1 - qi) / B;
      break;
    }
  }
  C = sqrt(Ci) * sqrt(Cj);
  if (A == 0 || C2 == 0) return 0;
  p.tau = -B; p.eu = A; p.ef = A * B; p.ru = C; p.rf = 6 * C;
  return orc_e6_cut(&p, ss, C2, corr);
}

/* Busing (sim/busing, POW = 0): u = f rho_ij exp((R_ij - r)/rho_ij) -
   C_ij/r^6, f = 0.05 e^2/A^2 = 0.05 electron^2 (sim/busing/sitesite.h:
   1-14, 112-115), the exp-6 macros with tau = iBij = -f/eps,
   ef = Aij = f exp(R_ij/rho_ij), eu = rhorho Aij, ru = C
   (initssaux(), sim/busing/sitesite.c:87-108); combrule()
   (sim/busing/sitesite.c:63-80): eps = EvdW_i + EvdW_j = f rho_ij, 0 when
   either is negative, which switches the pair off whole (SS_MEASURE tests
   eps, sim/busing/sitesite.h:248-256); R additive, C a product */
int oracle_setss_busing(mdx_exp6_t *ss, double Ei, double Ri, double Ci, double Ej,
                        double Rj, double Cj, int comb_rule, double C2, double *corr)
{
  const double fb = 0.05 * 408.779826922048 * 408.779826922048;
  double eps, rhorho;
  orc_e6 p;
  memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (comb_rule) return MDX_E_ARG;
  eps = (Ei < 0 || Ej < 0) ? 0 : Ei + Ej;
  if (eps == 0 || C2 == 0) return 0;
  rhorho = eps / fb;
  p.ef = fb * exp((Ri + Rj) / rhorho);
  p.tau = -fb / eps;
  p.eu = rhorho * p.ef;
  p.ru = Ci * Cj;
  p.rf = 6 * p.ru;
  return orc_e6_cut(&p, ss, C2, corr);
}

/* --------------------------------------------------------- erfc splines */

int oracle_erfc_table(mdx_erfc_t *tab, int *nout, double *sgridout, int grid,
                      double maxr, double cutoff, double alpha, int shift)
{
  int ito = (int)(sqr(maxr * alpha) * grid) + 1, n = ito, i;
  double h = 1.0 / grid, Du, Dd, x, xh, z, sgrid;
  double (*f)[2] = malloc(sizeof(double[2]) * (n + 3));
  if (!f) return MDX_E_ARG;
  for (i = n + 2; i >= 0; i--) {        /* erfcs(), SUBGRID=1 */
    if (i) {
      double rr = i * h, r = sqrt(rr), er = erfc(r);
      f[i][0] = er / r;
      f[i][1] = ((2 / SQRTPI) * exp(-rr) + er / r) / rr;
    } else f[i][0] = f[i][1] = 9e9;
  }
  Du = (-f[n + 2][0] + 8 * f[n + 1][0] - 8 * f[n - 1][0] + f[n - 2][0]) / (12 * h);
  Dd = (-f[n + 2][1] + 8 * f[n + 1][1] - 8 * f[n - 1][1] + f[n - 2][1]) / (12 * h);
  x = n * h;
  while (ito--) {
    mdx_erfc_t *p = tab + ito;
    double du, dd;
    xh = x;
    i = ito;
    x = i * h;
    du = f[i + 1][0] - f[i][0];
    p->Cu = (x * du - xh * Du * h) / (Du * h - du);
    p->Bu = -Du * sqr(p->Cu + xh);
    p->Au = f[i + 1][0] + Du * (p->Cu + xh);
    Du = -p->Bu / sqr(p->Cu + x);
    dd = f[i + 1][1] - f[i][1];
    p->Cd = (x * dd - xh * Dd * h) / (Dd * h - dd);
    p->Bd = -Dd * sqr(p->Cd + xh);
    p->Ad = f[i + 1][1] + Dd * (p->Cd + xh);
    Dd = -p->Bd / sqr(p->Cd + x);
    p->Au *= alpha; p->Ad *= alpha * alpha * alpha;
    p->Bu /= alpha; p->Bd *= alpha;
    p->Cu /= sqr(alpha); p->Cd /= sqr(alpha);
  }
  free(f);
  sgrid = grid * sqr(alpha);
  z = cutoff * cutoff;
  if (shift & 1) {
    mdx_erfc_t *p = tab + (int)(z * sgrid);
    double s = p->Bu / (p->Cu + z) + p->Au;
    for (i = 0; i < n; i++) tab[i].Au -= s;
  }
  if (shift & 2) {
    mdx_erfc_t *p = tab + (int)(z * sgrid);
    double s = p->Bd / (p->Cd + z) + p->Ad;
    for (i = 0; i < n; i++) tab[i].Ad -= s;
  }
  *nout = n;
  *sgridout = sgrid;
  return 0;
}

/* exacterud_sqrt(): -erf(alpha r)/r and its force/r (SUBGRID=1 branch) */
static double orc_exact(double rr, double alpha, double *erd)
{
  double alphaq = alpha * alpha, A = alpha * (-2 / SQRTPI), B = A * alphaq * 2;
  double err = alphaq * rr;
  if (err < 0.01) {
    int n = 1;
    double su = 0, sd = 0, q = 1, r;
    for (;;) {
      r = q / (2 * n + 1);
      sd += r * n;
      su += r;
      if (fabs(q) < 1e-16) break;
      q *= -err / ++n;
    }
    *erd = B * sd;
    return A * (1 - su * err);
  } else {
    double r = sqrt(rr), y = -erf(alpha * r) / r;
    *erd = (y - A * exp(-err)) / rr;
    return y;
  }
}

/* ------------------------------------------------------- Ewald set-up */

int oracle_ewald_autoset(double *alpha, double *kappa, double qmax, double sumq,
                         double cutoff, const double L[3], double epsr, double epsk)
{
  double V = L[0] * L[1] * L[2], a, k = *kappa, x;
  int nit = 10000;
  a = log(2 * qmax * sqrt(sumq / cutoff / V) / epsr);
  if (a < 0) return MDX_E_ARG;
  a = sqrt(a) / cutoff;
  do {
    if (!nit--) return MDX_E_NOCONV;
    x = k;
    k = log(qmax * a / (PI * epsk) * sqrt(8 * sumq / (k * V)));
    if (k <= 0) return MDX_E_ARG;
    k = a * sqrt(k) / PI;
  } while (fabs(x - k) > 1e-9);
  *alpha = a;
  *kappa = k;
  return 0;
}

static int orc_almostint(double x)
{
  x -= (int)x;
  return fabs(0.5 - fabs(x - 0.5)) < 1e-9;
}

int oracle_ewald_kvectors(double *kappa, const double L[3], int *nk)
{
  double dK = 2e-9;
  int nagain = 0;
again:
  if (nagain++ > 25) return MDX_E_NOCONV;
  {
    double Kk = sqr(*kappa);
    int kx, ky, kz, kxmax = (int)(L[0] * *kappa), iQ = 0;
    for (kx = 0; kx <= kxmax; kx++) {
      double kxq = Kk - sqr(kx / L[0]), aux = L[1] * sqrt(kxq);
      int kymax;
      if (orc_almostint(aux)) { *kappa += *kappa * dK; dK *= 1.4; goto again; }
      kymax = (int)aux;
      for (ky = 0; ky <= kymax; ky++) {
        double kyq = kxq - sqr(ky / L[1]);
        int kzmax;
        aux = L[2] * sqrt(kyq);
        if (orc_almostint(aux)) { *kappa += *kappa * dK; dK *= 1.4; goto again; }
        kzmax = (int)aux;
        for (kz = !(kx + ky); kz <= kzmax; kz++)
          iQ += 1 << ((kx > 0) + (ky > 0) + (kz > 0) - 1);
      }
    }
    *nk = iQ;
  }
  return 0;
}

static int orc_logint(double x)
{
  int i = x + 0.9999;
  if (i > 1)
    if (x / (i - 1) < i / x) i--;
  return i;
}

int oracle_lcsetup(int ncell[3], int ns, double percell, const double L[3])
{
  int i, prdc = 1, nundef = 0;
  double prdL = 1;
  if (percell <= 0) percell = 1 + log((double)ns);
  for (i = 0; i < 3; i++)
    if (ncell[i] == 0) { nundef++; prdL *= L[i]; }
    else prdc *= ncell[i];
  for (i = 0; i < 3; i++)
    if (ncell[i] == 0) {
      double x = pow(ns / (percell * prdc * prdL), 1. / nundef);
      ncell[i] = orc_logint(x * L[i]);
      if (ncell[i] < 2) ncell[i] = 2;
      prdL /= L[i];
      prdc *= ncell[i];
      nundef--;
    }
  return 0;
}

/* ------------------------------------------------------------- Ewald */

typedef struct { double re, im; } cplx;

typedef struct {
  int kxmax, *kymax, **kzmax;   /* ktab */
  int ntrip, nk;
} orc_kset;

/* the box the k-vector set and Ediag belong to: cook keeps the k-vectors
   chosen for the starting box when the box changes and recomputes only
   ekk with the new one (sim/ewald.c:570-613); Ediag is computed at the
   initialization only (sim/ewald.c:475-492).  NULL: the system's box.
   Set for the duration of one NPT step. */
static const double *orc_kbox = NULL;

static void orc_kset_build(orc_kset *K, double kappa, const double L[3])
{
  double Kk = kappa * kappa;
  int kx, ky, kz;
  K->kxmax = (int)(L[0] * kappa);
  K->kymax = malloc(sizeof(int) * (K->kxmax + 1));
  K->kzmax = malloc(sizeof(int *) * (K->kxmax + 1));
  K->ntrip = K->nk = 0;
  for (kx = 0; kx <= K->kxmax; kx++) {
    double kxq = Kk - sqr(kx / L[0]);
    K->kymax[kx] = (int)(L[1] * sqrt(kxq));
    K->kzmax[kx] = malloc(sizeof(int) * (K->kymax[kx] + 1));
    for (ky = 0; ky <= K->kymax[kx]; ky++) {
      double kyq = kxq - sqr(ky / L[1]);
      K->kzmax[kx][ky] = (int)(L[2] * sqrt(kyq));
      for (kz = !(kx + ky); kz <= K->kzmax[kx][ky]; kz++) {
        K->ntrip++;
        K->nk += 1 << ((kx > 0) + (ky > 0) + (kz > 0) - 1);
      }
    }
  }
}

static void orc_kset_free(orc_kset *K)
{
  int kx;
  for (kx = 0; kx <= K->kxmax; kx++) free(K->kzmax[kx]);
  free(K->kzmax);
  free(K->kymax);
}

#define CXE(A, E) { double c_ = (A).re; (A).re = c_ * (E).re - (A).im * (E).im; (A).im = c_ * (E).im + (A).im * (E).re; }
#define CXIE(A, E) { double c_ = (A).re; (A).re = c_ * (E).re + (A).im * (E).im; (A).im = (A).im * (E).re - c_ * (E).im; }

/* per charge recursion state: the qtab_t of sim/ewald.c:115-138 */
typedef struct { cplx x, xy, xY, xyz, xyZ, xYz, xYZ; } orc_qtab;

/* One thread's share [i0,i1) of the charges.  pass 1 sums Q(k) into Qk;
   pass 2 adds q*ek*Im(conj(Q) e) to F. Loop order and recursion as in
   sim/ewpass1.c and sim/ewpass2m.c. */
static void orc_ewald_pass(int pass, const orc_kset *K, const double *L,
                           int i0, int i1, const double *q, const cplx *ex,
                           const cplx *ey, const cplx *ez, orc_qtab *qt,
                           cplx *Qk, const double *ekk, double (*F)[3])
{
  int kx, ky, kz, iQ = 0, ikk = 0, i;
  for (i = i0; i < i1; i++) { qt[i].x.re = q[i]; qt[i].x.im = 0; }
  for (kx = 0; kx <= K->kxmax; kx++) {
    if (kx) for (i = i0; i < i1; i++) CXE(qt[i].x, ex[i]);
    for (i = i0; i < i1; i++) qt[i].xy = qt[i].xY = qt[i].x;
    for (ky = 0; ky <= K->kymax[kx]; ky++) {
      if (ky) for (i = i0; i < i1; i++) { CXE(qt[i].xy, ey[i]); CXIE(qt[i].xY, ey[i]); }
      for (i = i0; i < i1; i++) {
        qt[i].xyz = qt[i].xyZ = qt[i].xy;
        qt[i].xYz = qt[i].xYZ = qt[i].xY;
      }
      for (kz = !(kx + ky); kz <= K->kzmax[kx][ky]; kz++) {
        int ncase = (kx > 0) + (ky > 0) + (kz > 0);
        if (kz) for (i = i0; i < i1; i++) {
          CXE(qt[i].xyz, ez[i]); CXIE(qt[i].xyZ, ez[i]);
          CXE(qt[i].xYz, ez[i]); CXIE(qt[i].xYZ, ez[i]);
        }
        if (pass == 1) {
          cplx A = {0, 0}, B = {0, 0}, C = {0, 0}, D = {0, 0};
          if (ncase == 3) {
            for (i = i0; i < i1; i++) {
              A.re += qt[i].xyz.re; A.im += qt[i].xyz.im;
              B.re += qt[i].xyZ.re; B.im += qt[i].xyZ.im;
              C.re += qt[i].xYz.re; C.im += qt[i].xYz.im;
              D.re += qt[i].xYZ.re; D.im += qt[i].xYZ.im;
            }
            Qk[iQ++] = A; Qk[iQ++] = B; Qk[iQ++] = C; Qk[iQ++] = D;
          } else if (ncase == 2) {
            if (kz == 0)
              for (i = i0; i < i1; i++) {
                A.re += qt[i].xyz.re; A.im += qt[i].xyz.im;
                B.re += qt[i].xYz.re; B.im += qt[i].xYz.im;
              }
            else
              for (i = i0; i < i1; i++) {
                A.re += qt[i].xyz.re; A.im += qt[i].xyz.im;
                B.re += qt[i].xyZ.re; B.im += qt[i].xyZ.im;
              }
            Qk[iQ++] = A; Qk[iQ++] = B;
          } else {
            for (i = i0; i < i1; i++) { A.re += qt[i].xyz.re; A.im += qt[i].xyz.im; }
            Qk[iQ++] = A;
          }
        } else {
          double e = ekk[ikk], ek[3];
          ek[0] = e * (kx / L[0]); ek[1] = e * (ky / L[1]); ek[2] = e * (kz / L[2]);
          if (ncase == 3) {
            cplx A = Qk[iQ], B = Qk[iQ + 1], C = Qk[iQ + 2], D = Qk[iQ + 3];
            iQ += 4;
            for (i = i0; i < i1; i++) {
              double FA = A.re * qt[i].xyz.im - A.im * qt[i].xyz.re;
              double FB = B.re * qt[i].xyZ.im - B.im * qt[i].xyZ.re;
              double FC = C.re * qt[i].xYz.im - C.im * qt[i].xYz.re;
              double FD = D.re * qt[i].xYZ.im - D.im * qt[i].xYZ.re;
              F[i][0] += ek[0] * ((FA + FB) + (FC + FD));
              F[i][1] += ek[1] * ((FA + FB) - (FC + FD));
              F[i][2] += ek[2] * (FA - FB + FC - FD);
            }
          } else if (ncase == 2) {
            cplx A = Qk[iQ], B = Qk[iQ + 1];
            iQ += 2;
            if (kz) {
              for (i = i0; i < i1; i++) {
                double FA = A.re * qt[i].xyz.im - A.im * qt[i].xyz.re;
                double FB = B.re * qt[i].xyZ.im - B.im * qt[i].xyZ.re;
                if (kx) F[i][0] += ek[0] * (FA + FB);
                else F[i][1] += ek[1] * (FA + FB);
                F[i][2] += ek[2] * (FA - FB);
              }
            } else {
              for (i = i0; i < i1; i++) {
                double FA = A.re * qt[i].xyz.im - A.im * qt[i].xyz.re;
                double FB = B.re * qt[i].xYz.im - B.im * qt[i].xYz.re;
                F[i][0] += ek[0] * (FA + FB);
                F[i][1] += ek[1] * (FA - FB);
              }
            }
          } else {
            cplx A = Qk[iQ++];
            for (i = i0; i < i1; i++) {
              double FA = A.re * qt[i].xyz.im - A.im * qt[i].xyz.re;
              if (kx) F[i][0] += ek[0] * FA;
              else if (ky) F[i][1] += ek[1] * FA;
              else F[i][2] += ek[2] * FA;
            }
          }
        }
        ikk++;
      }
    }
  }
}

/* k-space part: returns En.el of Ewald(2) (measure) and adds forces of
   the charges to f[idx[i]]; energy/tensor bookkeeping of pass 2m */
/* Ediag of Ewald(0), sim/ewald.c:477-492: the k-vectors of the full
   lattice with kappa^2 < k^2 < kappa^2 + alpha^2, exp(-pi^2 k^2/alpha^2)/k^2
   each, times charges.sumq/(2 pi V) */
static double orc_ediag(const mdx_system_t *S, double sumq)
{
  const double *L = orc_kbox ? orc_kbox : S->L;
  double qkappax = sqr(S->kappa) + sqr(S->alpha), kappax = sqrt(qkappax), kk;
  double q = -sqr(PI / S->alpha), E = 0;
  int Kx = kappax * L[0], Ky = kappax * L[1], Kz = kappax * L[2], kx, ky, kz;
  for (kx = -Kx; kx <= Kx; kx++)
    for (ky = -Ky; ky <= Ky; ky++)
      for (kz = -Kz; kz <= Kz; kz++) {
        kk = sqr(kx / L[0]) + sqr(ky / L[1]) + sqr(kz / L[2]);
        if (kk > sqr(S->kappa) && kk < qkappax) E += exp(kk * q) / kk;
      }
  return E * sumq / (2 * PI * L[0] * L[1] * L[2]);
}

static double orc_ewald(const mdx_system_t *S, int nq, const int *idx,
                        const double *q, const double (*pos)[3], double (*f)[3],
                        int measure, double Pvir[6], int nthreads, double sumq_self)
{
  orc_kset K;
  const double *L = S->L;
  double V = L[0] * L[1] * L[2], PIaq = -sqr(PI / S->alpha);
  double *ekk, Energy = 0, PVIR[6] = {0, 0, 0, 0, 0, 0}, sumq = 0, M[3] = {0, 0, 0};
  double PIL[3] = {2 * PI / L[0], 2 * PI / L[1], 2 * PI / L[2]};
  cplx *ex, *ey, *ez, *Qk, *Qth;
  orc_qtab *qt;
  double (*F)[3];
  int i, kx, ky, kz, ikk, iQ, t;

  orc_kset_build(&K, S->kappa, orc_kbox ? orc_kbox : L);
  ekk = malloc(sizeof(double) * K.ntrip);
  ikk = 0;
  for (kx = 0; kx <= K.kxmax; kx++)
    for (ky = 0; ky <= K.kymax[kx]; ky++) {
      double aux = sqr(ky / L[1]) + sqr(kx / L[0]);
      for (kz = !(kx + ky); kz <= K.kzmax[kx][ky]; kz++) {
        double kk = sqr(kz / L[2]) + aux;
        ekk[ikk++] = 4 / (kk * (L[0] * L[1] * L[2])) * exp(PIaq * kk);
      }
    }
  ex = malloc(sizeof(cplx) * nq); ey = malloc(sizeof(cplx) * nq); ez = malloc(sizeof(cplx) * nq);
  qt = malloc(sizeof(orc_qtab) * nq);
  F = calloc(nq, sizeof(*F));
  Qk = calloc(K.nk, sizeof(cplx));
  Qth = calloc((size_t)K.nk * nthreads, sizeof(cplx));
  for (i = 0; i < nq; i++) {
    double a;
    M[0] += q[i] * pos[i][0]; M[1] += q[i] * pos[i][1]; M[2] += q[i] * pos[i][2];
    ex[i].re = cos(a = PIL[0] * pos[i][0]); ex[i].im = sin(a);
    ey[i].re = cos(a = PIL[1] * pos[i][1]); ey[i].im = sin(a);
    ez[i].re = cos(a = PIL[2] * pos[i][2]); ez[i].im = sin(a);
    sumq += q[i] * q[i];
  }
  /* pass 1: per-thread Q(k), summed serially (sim/ewald.c:756-780) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (t = 0; t < nthreads; t++)
    orc_ewald_pass(1, &K, L, (int)((long)nq * t / nthreads), (int)((long)nq * (t + 1) / nthreads),
                   q, ex, ey, ez, qt, Qth + (size_t)K.nk * t, ekk, F);
  for (t = 0; t < nthreads; t++)
    for (iQ = 0; iQ < K.nk; iQ++) {
      Qk[iQ].re += Qth[(size_t)K.nk * t + iQ].re;
      Qk[iQ].im += Qth[(size_t)K.nk * t + iQ].im;
    }
  /* pass 2: forces */
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (t = 0; t < nthreads; t++)
    orc_ewald_pass(2, &K, L, (int)((long)nq * t / nthreads), (int)((long)nq * (t + 1) / nthreads),
                   q, ex, ey, ez, qt, Qk, ekk, F);
  /* energy and tensor (ewpass2m.c:194-297) */
  if (measure) {
    ikk = iQ = 0;
    for (kx = 0; kx <= K.kxmax; kx++)
      for (ky = 0; ky <= K.kymax[kx]; ky++)
        for (kz = !(kx + ky); kz <= K.kzmax[kx][ky]; kz++) {
          double e = ekk[ikk++], kk, ptf, E1;
          int ncase = (kx > 0) + (ky > 0) + (kz > 0);
          double a2, b2 = 0, c2 = 0, d2 = 0, kxl = kx / L[0], kyl = ky / L[1], kzl = kz / L[2];
          kk = sqr(kxl) + sqr(kyl) + sqr(kzl);
          ptf = 1 / kk - PIaq;
          a2 = sqr(Qk[iQ].re) + sqr(Qk[iQ].im);
          if (ncase >= 2) b2 = sqr(Qk[iQ + 1].re) + sqr(Qk[iQ + 1].im);
          if (ncase == 3) {
            c2 = sqr(Qk[iQ + 2].re) + sqr(Qk[iQ + 2].im);
            d2 = sqr(Qk[iQ + 3].re) + sqr(Qk[iQ + 3].im);
          }
          iQ += 1 << (ncase - 1);
          Energy += E1 = (a2 + b2 + c2 + d2) * e;
          PVIR[0] += E1 * ptf * sqr(kxl);
          PVIR[1] += E1 * ptf * sqr(kyl);
          PVIR[2] += E1 * ptf * sqr(kzl);
          if (ncase == 3) {
            PVIR[3] += e * ptf * (a2 - b2 - c2 + d2) * kyl * kzl;
            PVIR[4] += e * ptf * (a2 - b2 + c2 - d2) * kxl * kzl;
            PVIR[5] += e * ptf * (a2 + b2 - c2 - d2) * kxl * kyl;
          } else if (ncase == 2) {
            if (kz) {
              PVIR[3] += e * ptf * (a2 - b2) * kyl * kzl;
              PVIR[4] += e * ptf * (a2 - b2) * kxl * kzl;
              PVIR[5] += e * ptf * (a2 + b2) * kxl * kyl;
            } else {
              PVIR[3] += e * ptf * (a2 - b2) * kyl * kzl;
              PVIR[4] += e * ptf * (a2 + b2) * kxl * kzl;
              PVIR[5] += e * ptf * (a2 - b2) * kxl * kyl;
            }
          } else {
            PVIR[3] += e * ptf * a2 * kyl * kzl;
            PVIR[4] += e * ptf * a2 * kxl * kzl;
            PVIR[5] += e * ptf * a2 * kxl * kyl;
          }
        }
    Energy /= 4 * PI;
    for (i = 0; i < 3; i++) Pvir[i] += Energy;
    for (i = 0; i < 6; i++) Pvir[i] -= 2 / (4 * PI) * PVIR[i];
  }
  /* cell-dipole term (sim/ewald.c:865-897): the Yeh-Berkowitz slab form
     in the el.corr directions (+8: no force) or the epsinf form; with el.bg
     and el.corr the charged-background field in those directions, added
     per charge with its virial as ADDFORCES of sim/ewald.c:901-931 does */
  {
    double c, ek[3] = {0, 0, 0}, cbg = 0, Edip = 0, qsum = 0;
    int k;
    for (i = 0; i < nq; i++) qsum += q[i];
    if (S->el_corr) {
      c = 4 * PI / V;
      for (k = 0; k < 3; k++)
        if (S->el_corr & 1 << k) {
          if (!(S->el_corr & 8)) ek[k] = -c * M[k];
          Edip += c / 2 * sqr(M[k]);
        }
    } else {
      c = 4 * PI / (2 * S->epsinf + 1) / V;
      ek[0] = -c * M[0]; ek[1] = -c * M[1]; ek[2] = -c * M[2];
      Edip = c / 2 * (sqr(M[0]) + sqr(M[1]) + sqr(M[2]));
    }
    if (measure) for (i = 0; i < 3; i++) Pvir[i] += Edip;
    if (S->el_bg && S->el_corr) cbg = 4 * PI * qsum / V;
    for (i = 0; i < nq; i++) {
      double *fi = f[idx[i]], ekx[3];
      for (k = 0; k < 3; k++) {
        ekx[k] = ek[k];
        if (cbg != 0 && (S->el_corr & 1 << k)) ekx[k] += cbg * pos[i][k];
      }
      fi[0] += F[i][0] + q[i] * ekx[0];
      fi[1] += F[i][1] + q[i] * ekx[1];
      fi[2] += F[i][2] + q[i] * ekx[2];
      if (measure) {
        Pvir[0] += q[i] * ekx[0] * pos[i][0];
        Pvir[1] += q[i] * ekx[1] * pos[i][1];
        Pvir[2] += q[i] * ekx[2] * pos[i][2];
        Pvir[3] += q[i] * ekx[1] * pos[i][2];
        Pvir[4] += q[i] * ekx[2] * pos[i][0];
        Pvir[5] += q[i] * ekx[0] * pos[i][1];
      }
    }
    free(ex); free(ey); free(ez); free(qt); free(F); free(Qk); free(Qth); free(ekk);
    orc_kset_free(&K);
    if (!measure) return 0;
    /* POLAR: sum of squared total site charges; the centre-Drude pair of a
       site is corrected in selffield() (sim/ewald.c:979-981, sim/scf.c:143) */
    if (sumq_self >= 0) sumq = sumq_self;
    /* the return of Ewald(2), sim/ewald.c:983-997: Ediag for el.diag = 1,
       the k-space energy, the background term of a net charge
       charges.bgE/V (bgE = -pi/2 (Q/alpha)^2, sim/ewaldprt.c:13), the self
       term and the dipole term */
    return (S->el_diag == 1 ? orc_ediag(S, sumq) : 0.0) + Energy
           - PI / 2 * sqr(qsum / S->alpha) / V
           - S->alphar / sqrt(PI) * sumq + Edip;
  }
}

/* -------------------------------------------------------- linked cells */

static double orc_minuseps(double x)
{
  double y = x * 0.999999999999999667, z;
  int i;
  for (i = 0; i < 5; i++) {
    volatile double zz = (x + y) / 2;
    z = zz;
    if (z >= x) return y;
    y = z;
  }
  return y;
}

typedef struct orc_site_s {
  double r[3], f[3];
  double rpol[3], fpol[3];   /* Drude offset and force (POLAR) */
  int i, n, st;
  double q, qpol;
  struct orc_site_s *next;
} orc_site;

typedef struct {
  const mdx_system_t *S;
  double U, Enel, Envir, Pvir[6];
  long long npair, nlj;
  int measure;
  /* RDF histograms of the measure path: records [nrec][nh] */
  unsigned long long *rdf;
  double rdf_grid;
  int rdf_nh, rdf_merged;
} orc_acc;

/* one pair of site types ti, tj at squared distance rr into the RDF
   records (cook/ljqqnp.c:218-225): bin (int)(r*grid), dropped from nhist
   on; the record of the type pair (i, j<=i) in the order initrdf() makes
   them (sim/simmeas.c:737-767), or the one merged record for grid<0 */
static void orc_rdf_count(orc_acc *a, int ti, int tj, double rr)
{
  int b = (int)(sqrt(rr) * a->rdf_grid), p = 0;
  if (b >= a->rdf_nh) return;
  if (!a->rdf_merged) {
    if (ti < tj) { int t = ti; ti = tj; tj = t; }
    p = ti * (ti + 1) / 2 + tj;
  }
  a->rdf[(size_t)p * a->rdf_nh + b]++;
}

static void orc_addf(orc_acc *a, double *f1, double *f2, double f, const double *dr)
{
  double fdr;
  f1[0] += fdr = f * dr[0]; f2[0] -= fdr;
  if (a->measure) { a->Pvir[0] += dr[0] * fdr; a->Pvir[5] += dr[1] * fdr; }
  f1[1] += fdr = f * dr[1]; f2[1] -= fdr;
  if (a->measure) { a->Pvir[1] += dr[1] * fdr; a->Pvir[3] += dr[2] * fdr; }
  f1[2] += fdr = f * dr[2]; f2[2] -= fdr;
  if (a->measure) { a->Pvir[2] += dr[2] * fdr; a->Pvir[4] += dr[0] * fdr; }
}

static double orc_erud(const mdx_system_t *S, double rr, double *erd)
{
  const mdx_erfc_t *p = S->erfc_tab + (int)(rr * S->erfc_sgrid);
  *erd = p->Bd / (p->Cd + rr) + p->Ad;
  return p->Bu / (p->Cu + rr) + p->Au;
}

/* site-site term of a non-excluded pair below C2: SS_MEASURE of the LJ
   form (sim/lj/sitesite.h:60-68) or of exp-6 (sim/exp6/sitesite.h:59-66)
   below C1, the quartic tail A (rr-C2q)^2 above (cook/ljqqnp.c:206-213);
   adds U and f(r)/r, the LJ virial and the LJ pair count */
static void orc_sitesite(orc_acc *a, int ti, int tj, double rr, double *U, double *f)
{
  const mdx_system_t *S = a->S;
  double C1q, C2q, A, A4;
  if (S->sstab_exp6) {
    const mdx_exp6_t *e = S->sstab_exp6 + ti * S->ntypes + tj;
    C1q = e->C1q; C2q = e->C2q; A = e->A; A4 = e->A4;
    if (rr >= C2q) return;
    if (rr < C1q) {
      double x = sqrt(rr), y = exp(e->tau * x), z = rr * rr * rr;
      *U = e->eu * y - e->ru / z;
      *f = e->ef * y / x - 6 * e->ru / z / rr;
    }
  } else {
    const mdx_lj_t *ss = S->sstab + ti * S->ntypes + tj;
    C1q = ss->C1q; C2q = ss->C2q; A = ss->A; A4 = ss->A4;
    if (rr >= C2q) return;
    if (rr < C1q) {
      double x = ss->Sq / rr, y = x * x * x, z = y * y;
      *U = ss->E4 * (z - y);
      *f = ss->E24 * (z + z - y) * x;
    }
  }
  if (rr >= C1q) {
    double x = rr - C2q;
    *U = x * x * A;
    *f = A4 * x;
  }
  a->Envir += rr * *f;
  a->nlj += A != 0;
}

/* SS_MEASURE_rep of the Buckingham and Busing forms (sim/buck/sitesite.h:
   145-157): the exponential part urep = eu e^(tau r) of the exp-6 table
   row of types ti, tj and frep = -urep'(r)/r; 0 from C2 on */
static double orc_rep(const mdx_system_t *S, int ti, int tj, double rr, double *frep)
{
  const mdx_exp6_t *e = S->sstab_exp6 + ti * S->ntypes + tj;
  double x, y;
  *frep = 0;
  if (rr >= e->C2q) return 0;
  x = sqrt(rr);
  y = exp(e->tau * x);
  *frep = e->ef * y / x;
  return e->eu * y;
}

/* a POLAR&1 shell-core term on a Drude-centre pair (cook/ljqqpol.c:
   561-592, 616-648): -kappa urep, force -kappa frep along d */
static void orc_shell(orc_acc *a, int ti, int tj, double kappa, double *f1, double *f2,
                      const double *d)
{
  double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], frep, urep;
  if (rr >= a->S->sstab_exp6[ti * a->S->ntypes + tj].C2q) return;
  urep = orc_rep(a->S, ti, tj, rr, &frep);
  frep *= kappa;
  a->Envir -= rr * frep;
  a->U -= kappa * urep;
  orc_addf(a, f1, f2, -frep, d);
}

/* polarLJqq[m] for POLAR=0 (cook/ljqqpol.c): the pair is chosen by the
   centre-centre distance; up to four charge-charge terms centre/Drude x
   centre/Drude, Drudes at r+rpol, LJ between the centres only.  With
   POLAR&1 shell-core data (shell_qkappa, shell_rep): anion1 = the central
   site has qkappa != 0 and the partner rep, anion2 the reverse
   (cook/ljqqpol.c:41-46); their terms follow cook/ljqqpol.c:525-535
   (centre-centre), 561-592 (Drude-centre) and 616-648 (centre-Drude). */
static void orc_ljqq_pol(orc_acc *a, orc_site *l, orc_site *ls)
{
  const mdx_system_t *S = a->S;
  double cq = S->cutoff * S->cutoff;
  for (; l; l = l->next) {
    double dr[3], d[3], rr, r2, f = 0, U = 0, qq, erd, u;
    double k1 = 0, k2 = 0;
    int k, an1 = 0, an2 = 0;
    if (S->shell_qkappa && S->sstab_exp6) {
      k1 = S->shell_qkappa[ls->st];
      k2 = S->shell_qkappa[l->st];
      an1 = k1 != 0 && S->shell_rep[l->st];
      an2 = k2 != 0 && S->shell_rep[ls->st];
    }
    for (k = 0; k < 3; k++) dr[k] = ls->r[k] - l->r[k];
    rr = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    if (rr >= cq) continue;
    a->npair++;
    if (ls->n == l->n) {
      int m0 = S->mol_first[ls->n];
      int ii = ls->i - m0, jj = l->i - m0;
      if ((S->excl[ls->i] >> jj & 1) || (S->excl[l->i] >> ii & 1)) {
        /* EXCL: exact Ewald corrections of all four charge pairs */
        if ((qq = ls->q * l->q) != 0) {
          u = orc_exact(rr, S->alphar, &erd);
          if (a->measure) a->Enel += u * qq;
          orc_addf(a, ls->f, l->f, erd * qq, dr);
        }
        if ((qq = ls->qpol * l->q) != 0) {
          for (k = 0; k < 3; k++) d[k] = dr[k] + ls->rpol[k];
          r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          u = orc_exact(r2, S->alphar, &erd);
          if (a->measure) a->Enel += u * qq;
          orc_addf(a, ls->fpol, l->f, erd * qq, d);
        }
        if (l->qpol != 0) {
          for (k = 0; k < 3; k++) d[k] = dr[k] - l->rpol[k];
          if (ls->q != 0) {
            qq = ls->q * l->qpol;
            r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            u = orc_exact(r2, S->alphar, &erd);
            if (a->measure) a->Enel += u * qq;
            orc_addf(a, ls->f, l->fpol, erd * qq, d);
          }
          if (ls->qpol != 0) {
            qq = ls->qpol * l->qpol;
            for (k = 0; k < 3; k++) d[k] += ls->rpol[k];
            r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            u = orc_exact(r2, S->alphar, &erd);
            if (a->measure) a->Enel += u * qq;
            orc_addf(a, ls->fpol, l->fpol, erd * qq, d);
          }
        }
        continue;
      }
    }
    /* RDF of the centres (cook/ljqqpol.c:477-485) */
    if (a->rdf) orc_rdf_count(a, ls->st, l->st, rr);
    orc_sitesite(a, ls->st, l->st, rr, &U, &f);
    a->U += U;
    if ((qq = ls->q * l->q) != 0) {
      if (an1 || an2) {
        double frep, urep = orc_rep(S, ls->st, l->st, rr, &frep);
        if (an1) { a->U += k1 * urep; a->Envir += k1 * frep * rr; f += k1 * frep; }
        if (an2) { a->U += k2 * urep; a->Envir += k2 * frep * rr; f += k2 * frep; }
      }
      u = orc_erud(S, rr, &erd);
      if (a->measure) a->Enel += u * qq;
      f += erd * qq;
    }
    orc_addf(a, ls->f, l->f, f, dr);
    if ((qq = ls->qpol * l->q) != 0) {
      for (k = 0; k < 3; k++) d[k] = dr[k] + ls->rpol[k];
      r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      u = orc_erud(S, r2, &erd);
      if (a->measure) a->Enel += u * qq;
      orc_addf(a, ls->fpol, l->f, erd * qq, d);
      if (an1) orc_shell(a, ls->st, l->st, k1, ls->fpol, l->f, d);
    }
    if (l->qpol != 0) {
      for (k = 0; k < 3; k++) d[k] = dr[k] - l->rpol[k];
      if (ls->q != 0) {
        qq = ls->q * l->qpol;
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        u = orc_erud(S, r2, &erd);
        if (a->measure) a->Enel += u * qq;
        orc_addf(a, ls->f, l->fpol, erd * qq, d);
        if (an2) orc_shell(a, ls->st, l->st, k2, ls->f, l->fpol, d);
      }
      if (ls->qpol != 0) {
        qq = ls->qpol * l->qpol;
        for (k = 0; k < 3; k++) d[k] += ls->rpol[k];
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        u = orc_erud(S, r2, &erd);
        if (a->measure) a->Enel += u * qq;
        orc_addf(a, ls->fpol, l->fpol, erd * qq, d);
      }
    }
  }
}

/* LJqq[m] for the central site ls against the list l (cook/ljqqnp.c) */
static void orc_ljqq(orc_acc *a, orc_site *l, orc_site *ls)
{
  const mdx_system_t *S = a->S;
  double cq = S->cutoff * S->cutoff;
  for (; l; l = l->next) {
    double dr[3], rr, f = 0, U = 0, qq, fdr;
    dr[0] = ls->r[0] - l->r[0];
    dr[1] = ls->r[1] - l->r[1];
    dr[2] = ls->r[2] - l->r[2];
    rr = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    if (rr >= cq) continue;
    a->npair++;
    if (ls->n == l->n) {
      int m0 = S->mol_first[ls->n];
      int ii = ls->i - m0, jj = l->i - m0;
      if ((S->excl[ls->i] >> jj & 1) || (S->excl[l->i] >> ii & 1)) {
        /* EXCL: exact Ewald erf correction (cook/ljqqnp.c:127-144) */
        if ((qq = ls->q * l->q) != 0) {
          double erd, u = orc_exact(rr, S->alphar, &erd);
          if (a->measure) a->Enel += u * qq;
          f = erd * qq;
        }
        goto addUf;
      }
    }
    orc_sitesite(a, ls->st, l->st, rr, &U, &f);
    if ((qq = ls->q * l->q) != 0) {
      const mdx_erfc_t *p = S->erfc_tab + (int)(rr * S->erfc_sgrid);
      double byerd = p->Bd / (p->Cd + rr) + p->Ad;
      double eru = p->Bu / (p->Cu + rr) + p->Au;
      if (a->measure) a->Enel += eru * qq;
      f += byerd * qq;
    }
    if (a->rdf) orc_rdf_count(a, ls->st, l->st, rr);
  addUf:
    a->U += U;
    ls->f[0] += fdr = f * dr[0]; l->f[0] -= fdr;
    if (a->measure) { a->Pvir[0] += dr[0] * fdr; a->Pvir[5] += dr[1] * fdr; }
    ls->f[1] += fdr = f * dr[1]; l->f[1] -= fdr;
    if (a->measure) { a->Pvir[1] += dr[1] * fdr; a->Pvir[3] += dr[2] * fdr; }
    ls->f[2] += fdr = f * dr[2]; l->f[2] -= fdr;
    if (a->measure) { a->Pvir[2] += dr[2] * fdr; a->Pvir[4] += dr[0] * fdr; }
  }
}

/* linklist() + serial linkcell(): half-shell traversal of cook/lc.c:217-460
   (periodic z, no SLIT).  Sites get cell ids from FP64 wrapped coordinates
   with the shaved rL of cook/linklist.c:81-86. Returns 0 or MDX_E_RANGE. */
static int orc_pairs(orc_acc *a, const double (*r)[3], double (*f)[3], int *cellid)
{
  const mdx_system_t *S = a->S;
  const int *nc = S->ncell;
  int ns = S->ns, i, n, k;
  double rL[3], cell[3], cq = S->cutoff * S->cutoff;
  orc_site *st = calloc(ns, sizeof(orc_site));
  orc_site **list = calloc((size_t)nc[0] * nc[1] * nc[2], sizeof(orc_site *));
  int ixs, iys, izs, err = 0;
#define LIST(X, Y, Z) list[((size_t)(X) * nc[1] + (Y)) * nc[2] + (Z)]
#define PAIRS(L_, LS_) (S->chargepol ? orc_ljqq_pol(a, L_, LS_) : orc_ljqq(a, L_, LS_))
  for (k = 0; k < 3; k++) {
    double x;
    volatile double t = nc[k] / S->L[k];
    rL[k] = t;
    x = orc_minuseps(S->L[k]);
    while ((int)(x * rL[k]) >= nc[k]) rL[k] = orc_minuseps(rL[k]);
  }
  for (n = 0; n < S->nmol; n++)
    for (i = S->mol_first[n]; i < S->mol_first[n + 1]; i++) {
      orc_site *l = st + i;
      int ic[3];
      for (k = 0; k < 3; k++) {
        l->r[k] = r[i][k];
        while (l->r[k] >= S->L[k]) l->r[k] -= S->L[k];
        ic[k] = (int)(l->r[k] * rL[k]);
        if (ic[k] >= nc[k] || ic[k] < 0) err = MDX_E_RANGE;
      }
      if (err) break;
      if (cellid) cellid[i] = (ic[0] * nc[1] + ic[1]) * nc[2] + ic[2];
      l->i = i; l->n = n; l->st = S->site_type[i]; l->q = S->charge[i];
      if (S->chargepol) {
        l->qpol = S->chargepol[i];
        for (k = 0; k < 3; k++) l->rpol[k] = r[ns + i][k];
      }
      l->next = LIST(ic[0], ic[1], ic[2]);
      LIST(ic[0], ic[1], ic[2]) = l;
    }
  if (err) { free(st); free(list); return err; }
  for (k = 0; k < 3; k++) { cell[k] = S->L[k] / nc[k]; rL[k] = nc[k] / S->L[k]; }

  for (ixs = 0; ixs < nc[0]; ixs++)
    for (iys = 0; iys < nc[1]; iys++)
      for (izs = 0; izs < nc[2]; izs++) {
        orc_site *ls;
        for (ls = LIST(ixs, iys, izs); ls; ls = ls->next) {
          orc_site lsloc = *ls;
          double cellr[3], DX, DY, dy, dz, dyq;
          int ix, ixm, ix1, iy, iym, iy0, iy1, iz, izm, iz0, iz1;
          for (k = 0; k < 3; k++) cellr[k] = rL[k] * lsloc.r[k];
          ix1 = cellr[0] + S->cutoff * rL[0];
          for (ix = ixm = ixs; ix <= ix1; ix++) {
            DX = ix == ixs ? 0 : ix - cellr[0];
            if (ixm == nc[0]) { lsloc.r[0] -= S->L[0]; ixm -= nc[0]; }
            dyq = cq - sqr(DX * cell[0]);
            if (dyq < 0) { if (dyq / cq < -0.1) err = MDX_E_RANGE; dyq = 0; }
            dy = sqrt(dyq) * rL[1];
            if (ixm == ixs) iy0 = iym = iys;
            else {
              iym = iy0 = (int)(cellr[1] + nc[1] - dy) - nc[1];
              if (iym < 0) { lsloc.r[1] += S->L[1]; iym += nc[1]; }
            }
            if (iym < 0 || iym > nc[1]) err = MDX_E_RANGE;
            iy1 = (int)(cellr[1] + dy);
            for (iy = iy0; iy <= iy1; iy++) {
              if (iym == nc[1]) { lsloc.r[1] -= S->L[1]; iym -= nc[1]; }
              if (iy > iys) DY = iy - cellr[1];
              else if (iy < iys) DY = iy + 1 - cellr[1];
              else DY = 0;
              dz = dyq - sqr(DY * cell[1]);
              if (dz < 0) { if (dz / cq < -0.1) err = MDX_E_RANGE; dz = 0; }
              dz = sqrt(dz) * rL[2];
              iz1 = (int)(cellr[2] + dz);
              if (ixm == ixs && iym == iys) {
                izm = iz0 = izs;
                for (iz = iz0; iz <= iz1; iz++) {
                  if (izm == nc[2]) { lsloc.r[2] -= S->L[2]; izm -= nc[2]; }
                  if (izm == izs) PAIRS(lsloc.next, &lsloc);
                  else PAIRS(LIST(ixm, iym, izm), &lsloc);
                  izm++;
                }
              } else {
                izm = iz0 = (int)(cellr[2] + nc[2] - dz) - nc[2];
                if (izm < 0) { lsloc.r[2] += S->L[2]; izm += nc[2]; }
                for (iz = iz0; iz <= iz1; iz++) {
                  if (izm == nc[2]) { lsloc.r[2] -= S->L[2]; izm -= nc[2]; }
                  PAIRS(LIST(ixm, iym, izm), &lsloc);
                  izm++;
                }
              }
              lsloc.r[2] = ls->r[2];
              iym++;
            }
            lsloc.r[1] = ls->r[1];
            ixm++;
          }
          for (k = 0; k < 3; k++) { ls->f[k] = lsloc.f[k]; ls->fpol[k] = lsloc.fpol[k]; }
        }
      }
  for (i = 0; i < ns; i++) for (k = 0; k < 3; k++) f[i][k] += st[i].f[k];
  if (S->chargepol)
    for (i = 0; i < ns; i++) for (k = 0; k < 3; k++) f[ns + i][k] += st[i].fpol[k];
#undef PAIRS
#undef LIST
  free(st);
  free(list);
  return err;
}

/* ------------------------------------------------------------ forces */

/* forces(B,A) of cook (cook/forces.c:180-203, serial): f is zeroed,
   Ewald(1), En.el=Ewald(2), pairforces() with En.el saved/restored. */
static int orc_forces(const mdx_system_t *S, const double *r, double *f, mdx_out_t *out,
                      int measure, int nthreads, int *cellid, orc_acc *rdf);

int oracle_forces(const mdx_system_t *S, const double *r, double *f,
                  mdx_out_t *out, int measure, int nthreads, int *cellid)
{
  return orc_forces(S, r, f, out, measure, nthreads, cellid, NULL);
}

/* a measured forces() call with the RDF histograms of its pairs added to
   hist (initrdf() records, sim/simmeas.c:655-767): grid is truncated to an
   integer as cook's abs() does, grid>0 one record per type pair (i, j<=i),
   grid<0 one record of all pairs, nh = (int)(cutoff*|grid|+1.0000001)
   bins each.  Returns MDX_E_ARG for grid 0 or nh out of [1, 100000]. */
int oracle_forces_rdf(const mdx_system_t *S, const double *r, double *f, mdx_out_t *out,
                      int nthreads, int *cellid, int grid, double cutoff,
                      unsigned long long *hist)
{
  orc_acc h;
  int ag = grid < 0 ? -grid : grid;
  memset(&h, 0, sizeof(h));
  if (!ag || !(cutoff >= 0 && cutoff <= 1e4)) return MDX_E_ARG;
  h.rdf_nh = (int)(cutoff * ag + 1.0000001);
  if (h.rdf_nh > 100000) return MDX_E_ARG;
  h.rdf = hist;
  h.rdf_grid = ag;
  h.rdf_merged = grid < 0;
  return orc_forces(S, r, f, out, 1, nthreads, cellid, &h);
}

static int orc_forces(const mdx_system_t *S, const double *r, double *f, mdx_out_t *out,
                      int measure, int nthreads, int *cellid, orc_acc *rdf)
{
  /* POLAR: r and f hold 2*ns rows, centres then Drude offsets/forces
     (polar_off = No.s*24, sim/siminit.c:640) */
  int ns = S->ns, i, nq = 0, err, npol = S->chargepol ? 2 : 1;
  int *idx = malloc(sizeof(int) * ns * npol);
  double *q = malloc(sizeof(double) * ns * npol);
  double (*pos)[3] = malloc(sizeof(double[3]) * ns * npol);
  double Pvir[6] = {0, 0, 0, 0, 0, 0}, Enel, sumq_self = -1;
  orc_acc a;
  if (nthreads < 1) nthreads = 1;
  memset(f, 0, sizeof(double) * 3 * ns * npol);
  for (i = 0; i < ns; i++) {
    if (S->charge[i] != 0) {
      idx[nq] = i; q[nq] = S->charge[i];
      pos[nq][0] = r[3 * i]; pos[nq][1] = r[3 * i + 1]; pos[nq][2] = r[3 * i + 2];
      nq++;
    }
    if (S->chargepol && S->chargepol[i] != 0) {
      int k;
      idx[nq] = ns + i; q[nq] = S->chargepol[i];
      for (k = 0; k < 3; k++) pos[nq][k] = r[3 * i + k] + r[3 * (ns + i) + k];
      nq++;
    }
  }
  if (S->chargepol) {
    sumq_self = 0;
    for (i = 0; i < ns; i++) sumq_self += sqr(S->charge[i] + S->chargepol[i]);
  }
  Enel = nq ? orc_ewald(S, nq, idx, q, (const double (*)[3])pos, (double (*)[3])f,
                        measure, Pvir, nthreads, sumq_self) : 0;
  memset(&a, 0, sizeof(a));
  if (rdf) a = *rdf;
  a.S = S;
  a.measure = measure;
  err = orc_pairs(&a, (const double (*)[3])r, (double (*)[3])f, cellid);
  memset(out, 0, sizeof(*out));
  out->En_kspace = Enel;
  out->En_el = a.Enel + Enel;
  out->En_pot = a.U;
  out->En_vir = a.Envir;
  for (i = 0; i < 6; i++) out->Pvir[i] = Pvir[i] + a.Pvir[i];
  out->npair = a.npair;
  out->nlj = a.nlj;
  free(idx); free(q); free(pos);
  return err;
}

/* exacterud_sqrt_1(): exacterud_sqrt() + 2 alpha/sqrt(pi), the Ewald
   correction of the centre-Drude pair of one site (sim/erexact.c:55-82) */
static double orc_exact1(double rr, double alpha, double *erd)
{
  double alphaq = alpha * alpha, A = alpha * (-2 / SQRTPI), B = A * alphaq * 2;
  double err = alphaq * rr;
  if (err < 0.01) {
    int n = 1;
    double su = 0, sd = 0, q = 1, r;
    for (;;) {
      r = q / (2 * n + 1);
      sd += r * n;
      su += r;
      if (fabs(q) < 1e-17) break;
      q *= -err / ++n;
    }
    *erd = B * sd;
    return A * (-su * err);
  } else {
    double r = sqrt(rr), y = -erf(alpha * r) / r;
    *erd = (y - A * exp(-err)) / rr;
    return y - A;
  }
}

/* One selffield() iteration (sim/scf.c:12-240, POLAR=0, no saturation or
   axial polarizability): centre-Drude Ewald correction, Drude move
   rpol <- (1-omega) rpol + omega alpha_qq f_pol, f_pol passed to the
   centre, self energy and Drude virial.  r, f: 2*ns rows.  Adds to
   out->En_el and out->Pvir; returns En.self in out->En_pot's place via
   *Eself, the largest |drpol|^2/omega^2*qpol^2 in *maxerr. */
int oracle_selffield(const mdx_system_t *S, double *r, double *f, mdx_out_t *out,
                     double omega, double *Eself, double *maxerr)
{
  int ns = S->ns, i, k;
  double self = 0, me = 0;
  if (!S->chargepol) return MDX_E_ARG;
  for (i = 0; i < ns; i++) {
    double *rp = r + 3 * (ns + i), *fp = f + 3 * (ns + i), *fc = f + 3 * i;
    double rr, ff, byerd, dd[3], e;
    if (S->chargepol[i] == 0) continue;
    rr = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
    if (rr >= 1.0) return MDX_E_RANGE;          /* el.minqq = 1 A */
    ff = S->charge[i] * S->chargepol[i];
    out->En_el += ff * orc_exact1(rr, S->alphar, &byerd);
    ff *= byerd;
    {
      mdx_out_t tmp;
      orc_acc a;
      memset(&a, 0, sizeof(a));
      a.measure = 1;
      orc_addf(&a, fp, fc, ff, rp);
      for (k = 0; k < 6; k++) out->Pvir[k] += a.Pvir[k];
      (void)tmp;
    }
    for (k = 0; k < 3; k++) {
      dd[k] = rp[k];
      rp[k] = (1 - omega) * rp[k] + omega * S->alpha_qq[i] * fp[k];
      dd[k] -= rp[k];
      fc[k] += fp[k];
    }
    self += rp[0] * fp[0] + rp[1] * fp[1] + rp[2] * fp[2];
    out->Pvir[0] -= rp[0] * fp[0];
    out->Pvir[1] -= rp[1] * fp[1];
    out->Pvir[2] -= rp[2] * fp[2];
    out->Pvir[3] -= rp[1] * fp[2];
    out->Pvir[4] -= rp[2] * fp[0];
    out->Pvir[5] -= rp[0] * fp[1];
    e = (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]) / (omega * omega) *
        sqr(S->chargepol[i]);
    if (e > me) me = e;
  }
  *Eself = self / 2;
  *maxerr = sqrt(me);
  return 0;
}

/* ------------------------------------------------------------- MD step */

/* One Shake() step (sim/constrd.c:622-1081) for SHAKE=1, VERLET=3,
   thermostat none, Berendsen or frictions (the Nose-Hoover and NPT steps
   below add their predictor through vpred, vfac, bscale).  r = cfg[0], r1 = cfg[1]
   are updated in place; f is scratch (cfg[2]).  Energies returned in out. */
static int orc_shake_step(const mdx_system_t *S, double *r, double *r1, const double *f,
                          const mdx_step_t *p, int measure, mdx_step_out_t *out,
                          const double *vpred, double vfac, double bscale);

int oracle_md_step(const mdx_system_t *S, double *r, double *r1, double *f,
                   const mdx_step_t *p, int measure, mdx_step_out_t *out,
                   int nthreads)
{
  int err;
  memset(out, 0, sizeof(*out));
  err = oracle_forces(S, r, f, &out->frc, measure, nthreads, NULL);
  if (err) return err;
  return orc_shake_step(S, r, r1, f, p, measure, out, NULL, 0, 1);
}

/* ASPC coefficients B_1..B_{kk+2} (sim/pred2.c:94-98) */
static void orc_aspc_coef(int kk, double *B)
{
  double o = (double)(4 * kk + 6) / (double)(kk + 3);
  int i;
  for (i = 0; i < kk + 2; i++) {
    B[i] = o * (i + 1);
    o *= (double)(i - kk - 1) / (double)(kk + i + 4);
  }
}

/* One Shake() step of a Drude system with ASPC (sim/pred2.c:25-99, option
   -p{k}20) followed by scf_iter selffield() iterations and the SHAKE step.
   r, r1, f have 2*ns rows.  As in cook, the Drude rows of r hold
   aa[0] = rpol(t) and the Drude rows of r1 hold aa[1] (sim/pred2.c:54);
   hist holds the older history aa[2..k+1] (k arrays of 3*ns, zero at the
   start like sim/pred2.c:52).  *kk is the predictor length reached so far
   (0 at the start; grows to k, sim/pred2.c:67-98).  scf_omega <= 0 takes
   the stability value (kk+2)/(2kk+3) (sim/pred2.c:75-77).  scf_iter < 0
   stops after the prediction (used to check the predictor alone).
   Returns En.self in out->En_self. */
int oracle_md_step_polar(const mdx_system_t *S, double *r, double *r1, double *f,
                         double *hist, int *kk, const mdx_step_t *p, int measure,
                         mdx_step_out_t *out, int nthreads)
{
  int ns = S->ns, n3 = 3 * ns, j, m, it, err, k = p->aspc_k;
  double B[12], omega = p->scf_omega;
  double *aa[12];
  memset(out, 0, sizeof(*out));
  if (k < 0 || k > 9 || *kk < 0 || *kk > k) return MDX_E_ARG;
  aa[0] = r + n3;
  aa[1] = r1 + n3;
  for (m = 2; m <= k + 1; m++) aa[m] = hist + (size_t)(m - 2) * n3;
  orc_aspc_coef(*kk, B);
  for (j = 0; j < n3; j++) {
    double pred = 0;
    for (m = 0; m < *kk + 2; m++) pred += B[m] * aa[m][j];
    for (m = k; m >= 0; m--) aa[m + 1][j] = aa[m][j];
    aa[0][j] = pred;
  }
  if (*kk < k) (*kk)++;
  if (omega <= 0) omega = (*kk + 2) / (double)(2 * *kk + 3);
  if (p->scf_iter < 0) return 0;
  for (it = 0; it < (p->scf_iter > 0 ? p->scf_iter : 1); it++) {
    double es, me;
    err = oracle_forces(S, r, f, &out->frc, measure, nthreads, NULL);
    if (err) return err;
    err = oracle_selffield(S, r, f, &out->frc, omega, &es, &me);
    if (err) return err;
    out->En_self = es;
    out->scf_maxerr = me;
  }
  return orc_shake_step(S, r, r1, f, p, measure, out, NULL, 0, 1);
}

/* blogT(): ln(E/T)/2 clamped to [-1, 1] (sim/constrd.c:98-107) */
static double orc_blogT(double E, double T)
{
  const double exp2 = 7.38905609893065022723;
  if (E > T * exp2) return 1;
  if (E * exp2 < T) return -1;
  return log(E / T) / 2;
}

/* bscale: the NPT constraint scaling of p_i - p_j (sim/constrd.c:820,
   the isotropic scale of sim/trvpscale.c:66-68; 1 otherwise) */
static int orc_shake_step(const mdx_system_t *S, double *r, double *r1, const double *f,
                          const mdx_step_t *p, int measure, mdx_step_out_t *out,
                          const double *vpred, double vfac, double bscale)
{
  int ns = S->ns, n, i, k, al;
  double hh = p->h * p->h, hscale = 0.5 / hh;
  double Ekin = 0, Pkin[6] = {0, 0, 0, 0, 0, 0}, virc = 0, PTvirc[6] = {0, 0, 0, 0, 0, 0};
  double Ktr = 0;
  double *pp = malloc(sizeof(double) * 3 * ns);
  int maxit0 = -16 * log(p->eps * 0.1), itmax = 0;
  long long nit = 0;                   /* constrit nit, sim/constrd.c:930 */
  for (n = 0; n < S->nmol; n++) {
    int i0 = S->mol_first[n], i1 = S->mol_first[n + 1];
    int c0 = S->cons_first[n], c1 = S->cons_first[n + 1], nc = c1 - c0, it = 0;
    int maxit = p->maxit > 0 ? p->maxit : maxit0 * (nc + 4);
    double omega = S->mol_omega ? S->mol_omega[n] : 0.6, maxsq;
    if (vpred)                         /* verlet.c, Nose with predicted velocities */
      for (i = i0; i < i1; i++)
        for (k = 0; k < 3; k++)
          pp[3 * i + k] = hh / S->mass[i] * f[3 * i + k] - vfac * vpred[3 * i + k]
                          + r[3 * i + k] + r1[3 * i + k];
    else                               /* verlet.c, no thermostat branch */
      for (i = i0; i < i1; i++)
        for (k = 0; k < 3; k++)
          pp[3 * i + k] = hh / S->mass[i] * f[3 * i + k] + r[3 * i + k] + r1[3 * i + k];
    if (nc) do {
      it++;
      maxsq = 0;
      for (al = c0; al < c1; al++) {
        int a = S->cons_ij[2 * al], b = S->cons_ij[2 * al + 1];
        double bondq = S->cons_bondq[al], sq = bondq, pij[3], rij[3];
        double mi = 1 / S->mass[a], mj = 1 / S->mass[b];
        for (k = 0; k < 3; k++) pij[k] = (pp[3 * a + k] - pp[3 * b + k]) * bscale;
        sq -= pij[0] * pij[0] + pij[1] * pij[1] + pij[2] * pij[2];
        if (fabs(sq / bondq) > maxsq) maxsq = fabs(sq / bondq);
        for (k = 0; k < 3; k++) rij[k] = r[3 * a + k] - r[3 * b + k];
        sq *= omega / (mi + mj);
        if (measure) {
          PTvirc[0] += (sq / bondq) * (rij[0] * rij[0]);
          PTvirc[1] += (sq / bondq) * (rij[1] * rij[1]);
          PTvirc[2] += (sq / bondq) * (rij[2] * rij[2]);
          PTvirc[3] += (sq / bondq) * (rij[1] * rij[2]);
          PTvirc[4] += (sq / bondq) * (rij[2] * rij[0]);
          PTvirc[5] += (sq / bondq) * (rij[0] * rij[1]);
        }
        virc += sq;
        sq /= bondq;                     /* SHAKE>0: rxp=bondq */
        for (k = 0; k < 3; k++) {
          rij[k] *= sq;
          pp[3 * a + k] += mi * rij[k];
          pp[3 * b + k] -= mj * rij[k];
        }
      }
    } while (maxsq > p->eps && it < maxit);
    if (nc && it == maxit && maxsq > p->eps) { free(pp); return MDX_E_NOCONV; }
    if (it > itmax) itmax = it;
    if (nc) nit += it;
    double vm1[3] = {0, 0, 0}, vm2[3] = {0, 0, 0}, mm = 0;
    for (i = i0; i < i1; i++) {        /* shakev3.c */
      double V1[3], V2[3], m = S->mass[i];
      for (k = 0; k < 3; k++) { V1[k] = r1[3 * i + k]; V2[k] = pp[3 * i + k] - r[3 * i + k]; }
      Ekin += (V1[0] * V1[0] + V1[1] * V1[1] + V1[2] * V1[2] +
               V2[0] * V2[0] + V2[1] * V2[1] + V2[2] * V2[2]) * m;
      Pkin[0] += (V1[0] * V1[0] + V2[0] * V2[0]) * m;
      Pkin[1] += (V1[1] * V1[1] + V2[1] * V2[1]) * m;
      Pkin[2] += (V1[2] * V1[2] + V2[2] * V2[2]) * m;
      Pkin[3] += (V1[1] * V1[2] + V2[1] * V2[2]) * m;
      Pkin[4] += (V1[2] * V1[0] + V2[2] * V2[0]) * m;
      Pkin[5] += (V1[0] * V1[1] + V2[0] * V2[1]) * m;
      for (k = 0; k < 3; k++) {
        vm1[k] += m * V1[k];
        vm2[k] += m * V2[k];
        r1[3 * i + k] = V2[k];
        r[3 * i + k] = pp[3 * i + k];
      }
      mm += m;
    }
    /* translational kinetic energy of the molecule (sim/shakev3.c:35) */
    Ktr += (vm1[0] * vm1[0] + vm1[1] * vm1[1] + vm1[2] * vm1[2] +
            vm2[0] * vm2[0] + vm2[1] * vm2[1] + vm2[2] * vm2[2]) / mm;
  }
  Ekin *= hscale;                          /* twice the kinetic energy */
  Ktr *= hscale;
  (void)ns;
  for (k = 0; k < 6; k++) { Pkin[k] *= hscale; PTvirc[k] /= hh; }
  virc /= hh;
  out->lambda = 1;
  if (p->thermostat == 1) {               /* Berendsen, sim/thermo.c:156-160 */
    double T = Ekin / p->nf, b;
    const double exp2 = 7.38905609893065022723;
    if (T > p->T * exp2) b = 1;
    else if (T * exp2 < p->T) b = -1;
    else b = log(T / p->T) / 2;
    out->lambda = exp(b * -p->h / p->tau);
    for (i = 0; i < 3 * ns; i++) r1[i] *= out->lambda;
  }
  if (p->thermostat >= 11 && p->thermostat <= 13) {
    /* decoupled friction thermostats (sim/thermo.c:107-135): ze on the
       molecular translations (T_TR = 11), z on the intramolecular motion
       (T_IN = 12), both with their own temperatures (T_FRICTIONS = 13) */
    double z = 0, ze = 0, f_tr = p->nf_tr, f_in = p->nf - p->nf_tr;
    if (p->thermostat == 11) ze = orc_blogT(Ktr / f_tr, p->T) * (p->h / p->tau);
    if (p->thermostat == 12) z = orc_blogT((Ekin - Ktr) / f_in, p->T) * (p->h / p->tau);
    if (p->thermostat == 13) {
      double T_tr = (f_in + f_tr) * p->T / (p->T_tr_in * f_in + f_tr);
      double T_in = (f_in + f_tr) * p->T / (f_in + f_tr / p->T_tr_in);
      z = orc_blogT((Ekin - Ktr) / f_in, T_in) * (p->h / p->tau);
      ze = orc_blogT(Ktr / f_tr, T_tr) * (p->h / p->tau);
    }
    ze -= z;
    for (n = 0; n < S->nmol; n++) {
      int i0 = S->mol_first[n], i1 = S->mol_first[n + 1];
      double vm[3] = {0, 0, 0}, mm = 0;
      for (i = i0; i < i1; i++) {
        for (k = 0; k < 3; k++) vm[k] += S->mass[i] * r1[3 * i + k];
        mm += S->mass[i];
      }
      for (k = 0; k < 3; k++) vm[k] /= mm;
      for (i = i0; i < i1; i++)
        for (k = 0; k < 3; k++) r1[3 * i + k] -= z * r1[3 * i + k] + ze * vm[k];
    }
  }
  out->En_kin_tr = Ktr / 2;
  out->En_kin = Ekin / 2;
  for (k = 0; k < 6; k++) { out->Pkin[k] = Pkin[k]; out->PTvirc[k] = PTvirc[k]; }
  out->virc = virc;
  out->shake_iter_max = itmax;
  out->shake_nit = nit;
  free(pp);
  return 0;
}

/* One Shake() step with the Nose-Hoover thermostat (thermostat == 2):
   the time-reversible velocity predictor of sim/trvpscale.c (length kk
   known so far, growing to p->trvp_k; hist[j-1] = level j of the history,
   [1 + 3 ns] doubles: logs velocity h dlogs/dt, then the h v of the
   sites), the Verlet step of sim/verlet.c with the friction -Vfactor
   vpred, SHAKE, and the leap-frog of logs of sim/thermo.c:180-200 (VERLET
   = 3 extended energy).  nose[0] = logs, nose[1] = logs(t) - logs(t-h)
   (cfg[1]->logs), updated; nose[2] = En.ext, nose[3] = En.dlogs out. */
int oracle_md_step_nose(const mdx_system_t *S, double *r, double *r1, double *f,
                        const mdx_step_t *p, double nose[4], int *kk, double *hist,
                        int measure, mdx_step_out_t *out, int nthreads)
{
  int ns = S->ns, stride = 1 + 3 * ns, k = *kk, j, i, err, knew;
  double b[10], vfac, hh = p->h * p->h, *vpred, ld, logs0, Tk;
  if (p->thermostat != 2 || p->trvp_k < 0 || p->trvp_k > 9 || k > p->trvp_k) return MDX_E_ARG;
  err = oracle_forces(S, r, f, &out->frc, measure, nthreads, NULL);
  if (err) return err;
  /* TRVP coefficients for k (sim/trvpscale.c:28-33) */
  b[0] = (double)(2 * k + 1) / (k + 1);
  for (j = 0; j < k; j++) b[j + 1] = -b[j] * (double)(k - j) / (k + j + 2);
  vpred = malloc(sizeof(double) * 3 * ns);
  vfac = b[0] * nose[1];
  for (j = 1; j <= k; j++) vfac += b[j] * hist[(size_t)(j - 1) * stride];
  for (i = 0; i < 3 * ns; i++) {
    double v = b[0] * r1[i];
    for (j = 1; j <= k; j++) v += b[j] * hist[(size_t)(j - 1) * stride + 1 + i];
    vpred[i] = v;
  }
  /* the history grows to trvp_k and shifts: level 1 gets (logs1, r1) */
  knew = k + 1 < p->trvp_k ? k + 1 : p->trvp_k;
  for (j = knew; j >= 2; j--)
    memcpy(hist + (size_t)(j - 1) * stride, hist + (size_t)(j - 2) * stride, sizeof(double) * stride);
  if (knew >= 1) {
    hist[0] = nose[1];
    memcpy(hist + 1, r1, sizeof(double) * 3 * ns);
  }
  *kk = knew;
  err = orc_shake_step(S, r, r1, f, p, measure, out, vpred, vfac, 1);
  free(vpred);
  if (err) return err;
  /* logs by leap-frog with the doubled kinetic energy (thermo.c) */
  Tk = 2 * out->En_kin / p->nf;
  ld = nose[1];
  logs0 = nose[0];
  nose[1] += hh * (Tk / p->T - 1) / (p->tau * p->tau);
  nose[0] += nose[1];
  nose[3] = (ld + nose[1]) / (2 * p->h);
  nose[2] = (logs0 + (p->tau / p->h) * (p->tau / p->h) * (ld * ld + nose[1] * nose[1]) / 4)
            * p->nf * p->T;
  return 0;
}

/* One Shake() step with the isotropic MTK (thermostat 22) or Hoover (23)
   barostat and the Nose thermostat, rescale="xyz": the predictor of
   sim/trvpscale.c:43-70 for logs, lambda and the sites (Vfactor = Vp0 +
   No.NPT Vp1, No.NPT = 1 + 3/No.f for MTK and 1 for Hoover, and the
   constraint scaling exp(2 lambda1 - lambda1(t-h)), or exp(lambda1) while
   the predictor is shorter than 2), the Verlet step of sim/verlet.c:22-36,
   SHAKE with the scaled p_i - p_j (sim/constrd.c:815-839), the pressure
   En.Pref of measurePconstraints() (sim/constrd.c:185-262: Ptr, the trace
   of the pressure tensor, or Pevir for virial = 1, + Ecorr/V^2), the
   leap-frog of logs and lambda of sim/thermo.c:58-99 with No.fx, No.M_T,
   No.M_P of sim/siminit.c:604-626 (No.ncoord = 3, No.Pkinq = 1), and the
   rescaling of the sites and the box by exp(lambda(t+h) - lambda(t))
   (rescalecfg(), sim/norm.c:989-1005).  L [3]: the box, updated; L0: the
   box of the k-vector set.  npt[0..2] = logs, logs1 (cfg[1]->logs),
   lambda1 (cfg[1]->lambda[0]), updated; npt[3] = En.ext, npt[4] = En.Pref,
   npt[5] = En.dlogs out.  hist [k][2 + 3 ns]: per level logs1, lambda1,
   then the h v of the sites; *kk as for Nose-Hoover.  No Drudes. */
int oracle_md_step_npt(const mdx_system_t *S, double *r, double *r1, double *f, double *L,
                       const double *L0, const mdx_step_t *p, double npt[6], int *kk,
                       double *hist, mdx_step_out_t *out, int nthreads)
{
  int ns = S->ns, stride = 2 + 3 * ns, k = *kk, j, i, err, knew;
  double b[10], hh = p->h * p->h, *vpred, Vp0 = 0, Vp1 = 0, vfac, bscale, NPT;
  double V, tr[3], Pref, kin, vir, lam0, scale, ldlogs, kinT, kinP, ext;
  double invf, fx, M_T, M_P, T = p->T;
  mdx_system_t SL = *S;
  if (!(p->thermostat == 22 || p->thermostat == 23) || p->trvp_k < 0 || p->trvp_k > 9 ||
      k > p->trvp_k || S->chargepol || !(p->tau_P > 0))
    return MDX_E_ARG;
  for (j = 0; j < 3; j++) SL.L[j] = L[j];
  orc_kbox = L0;
  err = oracle_forces(&SL, r, f, &out->frc, 1, nthreads, NULL);
  orc_kbox = NULL;
  if (err) return err;
  b[0] = (double)(2 * k + 1) / (k + 1);
  for (j = 0; j < k; j++) b[j + 1] = -b[j] * (double)(k - j) / (k + j + 2);
  for (j = 0; j <= k; j++) {
    Vp0 += b[j] * (j ? hist[(size_t)(j - 1) * stride] : npt[1]);
    Vp1 += b[j] * (j ? hist[(size_t)(j - 1) * stride + 1] : npt[2]);
  }
  NPT = 1. + (p->thermostat == 22) * 3. / p->nf;
  vfac = Vp0 + NPT * Vp1;
  bscale = k > 1 ? exp(2 * npt[2] - hist[1]) : exp(npt[2]);
  vpred = malloc(sizeof(double) * 3 * ns);
  for (i = 0; i < 3 * ns; i++) {
    double v = b[0] * r1[i];
    for (j = 1; j <= k; j++) v += b[j] * hist[(size_t)(j - 1) * stride + 2 + i];
    vpred[i] = v;
  }
  knew = k + 1 < p->trvp_k ? k + 1 : p->trvp_k;
  for (j = knew; j >= 2; j--)
    memcpy(hist + (size_t)(j - 1) * stride, hist + (size_t)(j - 2) * stride, sizeof(double) * stride);
  if (knew >= 1) {
    hist[0] = npt[1];
    hist[1] = npt[2];
    memcpy(hist + 2, r1, sizeof(double) * 3 * ns);
  }
  *kk = knew;
  err = orc_shake_step(&SL, r, r1, f, p, 1, out, vpred, vfac, bscale);
  free(vpred);
  if (err) return err;
  /* En.Pref at the volume of this step */
  V = L[0] * L[1] * L[2];
  kin = 2 * out->En_kin;                    /* En.kin, doubled */
  for (j = 0; j < 3; j++) tr[j] = (out->frc.Pvir[j] + out->PTvirc[j]) / V + out->Pkin[j] / V;
  vir = out->frc.En_vir + out->virc + out->frc.En_el;
  Pref = (p->virial == 1 ? (kin + vir) / 3 / V : (tr[0] + tr[1] + tr[2]) / 3) + p->Ecorr / (V * V);
  /* the leap-frog of logs and lambda (sim/thermo.c:58-99) */
  invf = (double)(p->thermostat == 22) / p->nf;
  fx = p->nf + (p->tau_P < .999e6 ? 1 : 0);
  M_T = p->tau * p->tau * T * p->nf;
  M_P = p->tau_P * p->tau_P * T * (p->nf + 3) * 3;
  ldlogs = npt[1];
  ext = fx * T * npt[0];
  kinT = M_T / 2 / hh * (Vp0 * Vp0);
  kinP = M_P / 2 / hh * (Vp1 * Vp1);
  npt[1] += hh * (kin + 6 * kinP - fx * T) / M_T;
  npt[0] += npt[1];
  npt[5] = (ldlogs + npt[1]) / (2 * p->h);
  ext += kinT + kinP * 3 + V * p->P;
  lam0 = log(L[0]);
  npt[2] += hh * (V * (Pref - p->P) + kin * invf) / M_P - Vp0 * Vp1;
  scale = exp((lam0 + npt[2]) - lam0);
  npt[3] = ext;
  npt[4] = Pref;
  for (i = 0; i < 3 * ns; i++) r[i] *= scale;
  for (j = 0; j < 3; j++) L[j] *= scale;
  return 0;
}

/* ------------------------------------------------- CPU baseline sample */

#include <time.h>
static double orc_now(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* Times the two halves of one force call on a bounded sample: the full
   r-space pass (linked cells + pair kernel) and the k-space passes over
   every kstride-th charge only (the k-space cost is linear in the number
   of charges for a fixed k-set).  t[0] = r-space seconds, t[1] = sampled
   k-space seconds, t[2] = charges in the sample. */
int oracle_time_sample(const mdx_system_t *S, const double *r, int nthreads,
                       int kstride, double t[3])
{
  /* r: [ns] rows, or [2 ns] (centres, Drude offsets) with Drudes; the
     k-space sample takes every kstride-th charge of the charge list
     (centres, then Drudes at r + rpol) */
  int ns = S->ns, i, nq = 0, err, polar = S->chargepol != NULL;
  int nrow = polar ? 2 * ns : ns;
  int *idx = malloc(sizeof(int) * nrow);
  double *q = malloc(sizeof(double) * nrow);
  double (*pos)[3] = malloc(sizeof(double[3]) * nrow);
  double *f = calloc(3 * (size_t)nrow, sizeof(double));
  double Pvir[6] = {0, 0, 0, 0, 0, 0}, t0;
  orc_acc a;
  int k = 0;
  if (kstride < 1) kstride = 1;
  for (i = 0; i < nrow; i++) {
    int site = i < ns ? i : i - ns, d;
    double qi = i < ns ? S->charge[i] : S->chargepol[site];
    if (qi == 0 || (k++ % kstride) != 0) continue;
    idx[nq] = i;
    q[nq] = qi;
    for (d = 0; d < 3; d++) pos[nq][d] = r[3 * site + d] + (i < ns ? 0 : r[3 * i + d]);
    nq++;
  }
  memset(&a, 0, sizeof(a));
  a.S = S;
  a.measure = 0;
  t0 = orc_now();
  err = orc_pairs(&a, (const double (*)[3])r, (double (*)[3])f, NULL);
  t[0] = orc_now() - t0;
  t0 = orc_now();
  if (nq) orc_ewald(S, nq, idx, q, (const double (*)[3])pos, (double (*)[3])f, 0, Pvir, nthreads, -1);
  t[1] = orc_now() - t0;
  t[2] = nq;
  free(idx); free(q); free(pos); free(f);
  return err;
}
