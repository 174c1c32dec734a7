// Host-side set-up of the force path: the numbers cook computes once at
// start-up and hands to forces() through its globals.  Each helper states
// the reference routine whose arithmetic it reproduces; the formulas have
// to be the reference's (they define the model), the code is this
// project's own.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "mdx_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.772453850905516;
// the electron charge in program units, sqrt(K A) (sim/units.h:108)
constexpr double kElectron = 408.779826922048;
// No.eps = machprec() (sim/siminit.c:64-66), the convergence scale of setss
constexpr double kMachEps = 1.1102230246251565e-16;

inline double p6(double x) { double y = x * x * x; return y * y; }

// LJ in van der Waals form u = -eps[(sig/r)^12 - 2(sig/r)^6], eps < 0
// (sim/lj/sitesite.h:79-80)
struct VdW {
  double sig, eps;
  double u(double r) const { double x = p6(sig / r); return -eps * (x * x - 2 * x); }
  double f(double r) const { double x = p6(sig / r); return -12 * eps * (x * x - x) / r; }
};

// combrule() for the square-root energy rule (comb_rule&1, also chosen by
// cook when a site has no polarizability), sim/lj/sitesite.c:66-93
VdW combine(double Ei, double Ri, double Ej, double Rj, int comb_rule) {
  double r = (comb_rule & 2) ? std::pow(4 * Ri * Rj, 3) : p6(Ri + Rj);
  double B = std::sqrt(-2 * Ei) * std::sqrt(-2 * Ej) * r;
  double A = B * r;
  VdW v;
  v.sig = (B == 0) ? 0 : std::pow(A / B, 1.0 / 6);
  v.eps = (A == 0) ? 0 : B * B / (-2 * A);
  return v;
}

int logint(double x) {  // cook/linklist.c:294-302
  int i = (int)(x + 0.9999);
  if (i > 1 && x / (i - 1) < i / x) i--;
  return i;
}

bool almost_integer(double x) {  // sim/ewald.c:240-250, machineprec=1e-9
  x -= (int)x;
  return std::fabs(0.5 - std::fabs(x - 0.5)) < 1e-9;
}

// exp-6 of sim/exp6/sitesite.h:59-75 with the constants of initssaux()
// (sim/exp6/sitesite.c:78-85): u = eu e^(tau r) - ru/r^6,
// f(r)/r = ef e^(tau r)/r - 6 ru/r^8; eps > 0 is the well depth |E_min|
// at r = sig, a the stiffness
struct Exp6 {
  double tau, eu, ef, ru, rf;
  Exp6() = default;
  Exp6(double sig, double eps, double a) {
    tau = -(a / sig);
    eu = 6 * eps / (a - 6) * std::exp(a);
    ef = eu * (a / sig);
    ru = eps * a * p6(sig) / (a - 6);
    rf = 6 * ru;
  }
  // ss_u() and ss_f() of sim/setss.c:105-134: the macro evaluated from
  // rr = r^2, U and f*r (the force, f being force/r)
  double u(double r) const {
    double rr = r * r, x = std::sqrt(rr);
    return eu * std::exp(tau * x) - ru / (rr * rr * rr);
  }
  double f(double r) const {
    double rr = r * r, x = std::sqrt(rr), z = rr * rr * rr;
    return (ef * std::exp(tau * x) / x - rf / z / rr) * r;
  }
};

// smooth cutoff of setss() (sim/setss.c:244-306), potential-independent:
// the fixed point for C1 such that the quartic A (r^2 - C2^2)^2 on
// [C1,C2] matches u and f at C1; then the homogeneous tail correction
// (sim/setss.c:455-499): the analytic quartic part + 2-point Gauss in 1/r
// on [C1,inf) with NGauss4 = 64 intervals (sim/setss.c:49)
template <class Form>
int smooth_cutoff(const Form &v, double C2, const char *who, double &C1q, double &C2q,
                  double &A, double &A4, double *corr) {
  double C1 = 0.775 * C2, prev;
  int n = 0;
  do {
    double fc = v.f(C1);
    if (std::fabs(fc) < 1e-50) return mdx::fail(MDX_E_ARG, std::string(who) + ": zero force at C1");
    prev = C1;
    double t = 1 + 4 * v.u(C1) / C1 / fc;
    if (n++ > 2000 || t < 0)
      return mdx::fail(MDX_E_NOCONV, std::string(who) +
                                         ": no consistent smooth cutoff (LJcutoff too short)");
    C1 = C2 / std::sqrt(t);
  } while (std::fabs(1 - C1 / prev) > kMachEps * n);
  if (C1 > 0.999 * C2 || C1 < 0.5 * C2) return mdx::fail(MDX_E_ARG, std::string(who) + ": bad cutoff");
  double d = C2 * C2 - C1 * C1;
  A = v.u(C1) / (d * d);
  A4 = -v.f(C1) / d / C1;
  C1q = C1 * C1;
  C2q = C2 * C2;
  if (corr) {
    const int ng = 64;
    const double g = 0.2113248654051871;
    double c0 = 4 * kPi * A *
                (C1 * (p6(C1) / 7 - 0.4 * std::pow(C1, 4) * C2 * C2 + C1 * C1 * std::pow(C2, 4) / 3) -
                 8.0 / 105 * std::pow(C2, 7));
    double h = 1.0 / ng / C1, c1 = 0;
    for (int k = 0; k < ng; k++) {
      double r1 = 1 / ((k + g) * h), r2 = 1 / ((k + 1 - g) * h);
      c1 += std::pow(r1, 4) * v.u(r1) + std::pow(r2, 4) * v.u(r2);
    }
    *corr = c0 + 2 * kPi * h * c1;
  }
  return 0;
}

}  // namespace

extern "C" int mdx_setss(mdx_lj_t *ss, double Ei, double Ri, double Ej,
                         double Rj, int comb_rule, double C2, double *corr) {
  if (!ss) return mdx::fail(MDX_E_ARG, "mdx_setss: ss is NULL");
  std::memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (Ei > 0 || Ej > 0) return mdx::fail(MDX_E_ARG, "mdx_setss: EvdW must be <= 0");
  VdW v = combine(Ei, Ri, Ej, Rj, comb_rule);
  // initssaux(), sim/lj/sitesite.c:96-104: LJsigma^2 = sig^2 / 2^(1/3)
  ss->Sq = v.sig * v.sig * 0.7937005259840998;
  ss->E4 = -4 * v.eps;
  if (ss->Sq != 0) {
    ss->E24 = -24 * v.eps / ss->Sq;
    ss->E48 = -48 * v.eps / ss->Sq;
  }
  if (v.eps == 0 || C2 == 0) return 0;  // no LJ term: C1=C2=A=A4=0
  return smooth_cutoff(v, C2, "mdx_setss", ss->C1q, ss->C2q, ss->A, ss->A4, corr);
}

extern "C" int mdx_setss_exp6(mdx_exp6_t *ss, double Ei, double Ri, double ai, double Ej,
                              double Rj, double aj, int comb_rule, double C2, double *corr) {
  if (!ss) return mdx::fail(MDX_E_ARG, "mdx_setss_exp6: ss is NULL");
  std::memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (Ei > 0 || Ej > 0) return mdx::fail(MDX_E_ARG, "mdx_setss_exp6: EvdW must be <= 0");
  if (!(comb_rule & 1))
    return mdx::fail(MDX_E_ARG, "mdx_setss_exp6: comb_rule&1 = 0 (Kirkwood) is not implemented");
  // combrule(), sim/exp6/sitesite.c:60-76: geometric (2) or additive
  // diameters, geometric mean of the (negative) minima and of a
  const double sig = (comb_rule & 2) ? std::sqrt(Ri * Rj * 4) : Ri + Rj;
  const double eps = std::sqrt(Ei * Ej), a = std::sqrt(ai * aj);
  if (eps == 0 || C2 == 0) return 0;    // no exp-6 term: the table is zero
  if (!(sig > 0) || !(a > 6))
    return mdx::fail(MDX_E_ARG, "mdx_setss_exp6: need sigma > 0 and a > 6");
  Exp6 v(sig, eps, a);
  ss->tau = v.tau;
  ss->eu = v.eu;
  ss->ef = v.ef;
  ss->ru = v.ru;
  return smooth_cutoff(v, C2, "mdx_setss_exp6", ss->C1q, ss->C2q, ss->A, ss->A4, corr);
}

extern "C" int mdx_setss_buck(mdx_exp6_t *ss, double Ai, double Bi, double Ci, double Aj,
                              double Bj, double Cj, int comb_rule, double C2, double *corr) {
  if (!ss) return mdx::fail(MDX_E_ARG, "mdx_setss_buck: ss is NULL");
  std::memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (((comb_rule & 12) == 0) != ((comb_rule & 3) == 0))
    return mdx::fail(MDX_E_ARG, "mdx_setss_buck: Waldman-Hagler must be set for both A and B or for "
                                "none (sim/buck/sitesite.c:85-87)");
  if (Ai < 0 || Aj < 0 || Bi < 0 || Bj < 0 || Ci < 0 || Cj < 0)
    return mdx::fail(MDX_E_ARG, "mdx_setss_buck: A, B, C must be >= 0");
  // combrule() of sim/buck/sitesite.c:113-170 (POW = 0)
  double B = 0, A = 0, aux = 0;
  const double Bi6 = 0.5 / p6(Bi), Bj6 = 0.5 / p6(Bj);
  if ((comb_rule & 3) == 2 || Bi * Bj != 0) switch (comb_rule & 3) {
    case 0: aux = Bi6 + Bj6; B = std::pow(aux, -1. / 6); break;
    case 1: B = std::sqrt(Bi) * std::sqrt(Bj); break;
    case 2: B = (Bi + Bj) / 2; break;
    case 3: B = 1 / (0.5 / Bi + 0.5 / Bj); break;
  }
  if (Ai * Aj != 0) switch (comb_rule & 12) {
    case 0: A = std::sqrt(Ai) / (Bi * Bi * Bi) * (std::sqrt(Aj) / (Bj * Bj * Bj)) / aux; break;
    case 4: A = std::sqrt(Ai) * std::sqrt(Aj); break;
    case 8: {                                     // Kong
      const double qi = Bi / (Bi + Bj), x = (Aj * Bj) / (Ai * Bi);
      A = (Ai * std::pow(x, qi) + Aj * std::pow(x, qi - 1)) / 2;
      break;
    }
    case 12: {                                    // Busing
      const double qi = (0.5 / Bi) / (0.5 / Bi + 0.5 / Bj);
      A = std::pow(Ai * Bi, qi) * std::pow(Aj * Bj, 1 - qi) / B;
      break;
    }
  }
  const double C = std::sqrt(Ci) * std::sqrt(Cj);
  if (A == 0 || C2 == 0) return 0;                // no site-site term: the table is zero
  // initssaux(), sim/buck/sitesite.c:178-188: the exp-6 form's constants
  Exp6 v;
  v.tau = -B;
  v.eu = A;
  v.ef = A * B;
  v.ru = C;
  v.rf = 6 * C;
  ss->tau = v.tau;
  ss->eu = v.eu;
  ss->ef = v.ef;
  ss->ru = v.ru;
  return smooth_cutoff(v, C2, "mdx_setss_buck", ss->C1q, ss->C2q, ss->A, ss->A4, corr);
}

extern "C" int mdx_setss_busing(mdx_exp6_t *ss, double Ei, double Ri, double Ci, double Ej,
                                double Rj, double Cj, int comb_rule, double C2, double *corr) {
  if (!ss) return mdx::fail(MDX_E_ARG, "mdx_setss_busing: ss is NULL");
  std::memset(ss, 0, sizeof(*ss));
  if (corr) *corr = 0;
  if (comb_rule != 0)
    return mdx::fail(MDX_E_ARG, "mdx_setss_busing: the Busing combining rule is comb_rule=0 "
                                "(sim/busing/sitesite.c:37)");
  // combrule() of sim/busing/sitesite.c:63-80: rho and R additive, C a
  // product; a site of EvdW < 0 switches every pair of it off
  const double eps = (Ei < 0 || Ej < 0) ? 0.0 : Ei + Ej;
  const double R = Ri + Rj, C = Ci * Cj;
  if (eps == 0 || C2 == 0) return 0;    // SS_MEASURE skips eps = 0 whole: the table is zero
  // initssaux(), sim/busing/sitesite.c:87-108: rho = eps/f,
  // u_rep = rho A e^(-r/rho) with A = f e^(R/rho), f = 0.05 e^2/A^2
  const double f = 0.05 * kElectron * kElectron, rho = eps / f;
  Exp6 v;
  v.tau = -f / eps;
  v.ef = f * std::exp(R / rho);
  v.eu = rho * v.ef;
  v.ru = C;
  v.rf = 6 * C;
  ss->tau = v.tau;
  ss->eu = v.eu;
  ss->ef = v.ef;
  ss->ru = v.ru;
  return smooth_cutoff(v, C2, "mdx_setss_busing", ss->C1q, ss->C2q, ss->A, ss->A4, corr);
}

extern "C" int mdx_erfc_size(int grid, double maxr, double alpha) {
  return (int)(maxr * alpha * maxr * alpha * grid) + 1;
}

// Hyperbolic splines A + B/(C+x), x = r^2, of eru = erfc(alpha r)/r and
// erd = -(1/r) d eru/dr on the grid 1/grid in (alpha r)^2, built from the
// top down so that value and slope are continuous (sim/splines.c:23-100,
// SPLINE=-2, SUBGRID=1 with library erfc, sim/erfc.c:68-80); shift&1/&2
// subtract eru/erd at the cutoff (sim/erfc.c:177-187).
extern "C" int mdx_erfc_table(mdx_erfc_t *tab, int *nout, double *sgrid,
                              int grid, double maxr, double cutoff,
                              double alpha, int shift) {
  if (!tab || !nout || !sgrid || grid <= 0 || alpha <= 0 || maxr < cutoff)
    return mdx::fail(MDX_E_ARG, "mdx_erfc_table: bad arguments");
  const int n = mdx_erfc_size(grid, maxr, alpha);
  const double h = 1.0 / grid;
  std::vector<double> u(n + 3), d(n + 3);
  for (int i = n + 2; i >= 0; i--) {
    if (i == 0) { u[0] = d[0] = 9e9; continue; }
    double x = i * h, r = std::sqrt(x), e = std::erfc(r);
    u[i] = e / r;
    d[i] = ((2 / kSqrtPi) * std::exp(-x) + e / r) / x;
  }
  // 4th-order one-sided start of the slope at the top node
  double su = (-u[n + 2] + 8 * u[n + 1] - 8 * u[n - 1] + u[n - 2]) / (12 * h);
  double sd = (-d[n + 2] + 8 * d[n + 1] - 8 * d[n - 1] + d[n - 2]) / (12 * h);
  const double a1 = alpha, a2 = alpha * alpha, a3 = a2 * alpha;
  double xhi = n * h;
  for (int i = n - 1; i >= 0; i--) {
    double xlo = i * h;
    mdx_erfc_t &p = tab[i];
    // fit B/(C+x)+A through (xlo,y[i]), (xhi,y[i+1]) with slope s at xhi
    auto fit = [&](const std::vector<double> &y, double &s, double &A, double &B, double &C) {
      double dy = y[i + 1] - y[i];
      C = (xlo * dy - xhi * s * h) / (s * h - dy);
      B = -s * (C + xhi) * (C + xhi);
      A = y[i + 1] + s * (C + xhi);
      s = -B / ((C + xlo) * (C + xlo));  // slope at the lower node
    };
    fit(u, su, p.Au, p.Bu, p.Cu);
    fit(d, sd, p.Ad, p.Bd, p.Cd);
    // rescale from the alpha=1 variable (alpha r)^2 to r^2
    p.Au *= a1; p.Bu /= a1; p.Cu /= a2;
    p.Ad *= a3; p.Bd *= a1; p.Cd /= a2;
    xhi = xlo;
  }
  const double sg = grid * a2, z = cutoff * cutoff;
  const int iz = (int)(z * sg);
  if (iz >= n) return mdx::fail(MDX_E_ARG, "mdx_erfc_table: cutoff outside the table");
  if (shift & 1) {
    double s = tab[iz].Bu / (tab[iz].Cu + z) + tab[iz].Au;
    for (int i = 0; i < n; i++) tab[i].Au -= s;
  }
  if (shift & 2) {
    double s = tab[iz].Bd / (tab[iz].Cd + z) + tab[iz].Ad;
    for (int i = 0; i < n; i++) tab[i].Ad -= s;
  }
  *nout = n;
  *sgrid = sg;
  return 0;
}

// Ewaldparm(): alpha from the r-space error target, kappa by fixed-point
// iteration from the k-space target (random-charge estimates),
// sim/ewald.c:1044-1057.
extern "C" int mdx_ewald_autoset(double *alpha, double *kappa, double qmax,
                                 double sumq, double cutoff, const double L[3],
                                 double epsr, double epsk) {
  if (!alpha || !kappa || sumq <= 0 || cutoff <= 0)
    return mdx::fail(MDX_E_ARG, "mdx_ewald_autoset: bad arguments");
  const double V = L[0] * L[1] * L[2];
  double a = std::log(2 * qmax * std::sqrt(sumq / cutoff / V) / epsr);
  if (a < 0) return mdx::fail(MDX_E_ARG, "mdx_ewald_autoset: cannot determine alpha (epsr)");
  a = std::sqrt(a) / cutoff;
  double k = *kappa > 0 ? *kappa : 0.2, prev;
  int left = 10000;
  do {
    if (!left--) return mdx::fail(MDX_E_NOCONV, "mdx_ewald_autoset: kappa iteration");
    prev = k;
    double t = std::log(qmax * a / (kPi * epsk) * std::sqrt(8 * sumq / (k * V)));
    if (t <= 0) return mdx::fail(MDX_E_ARG, "mdx_ewald_autoset: cannot determine kappa (epsk)");
    k = a * std::sqrt(t) / kPi;
  } while (std::fabs(prev - k) > 1e-9);
  *alpha = a;
  *kappa = k;
  return 0;
}

// k-vector census of Ewald(0) with the fix for limits that are almost
// integers (kappa raised by 2e-9 relative, growing by 1.4 each retry),
// sim/ewald.c:484-522.  Counts each (kx,ky,kz) with its sign variants.
extern "C" int mdx_ewald_kvectors(double *kappa, const double L[3], int *nk) {
  if (!kappa || !nk || *kappa <= 0) return mdx::fail(MDX_E_ARG, "mdx_ewald_kvectors: bad kappa");
  double dK = 2e-9;
  for (int attempt = 0; attempt <= 25; attempt++) {
    const double K2 = *kappa * *kappa;
    const int kxmax = (int)(L[0] * *kappa);
    long count = 0;
    bool retry = false;
    for (int kx = 0; kx <= kxmax && !retry; kx++) {
      double rx = K2 - (kx / L[0]) * (kx / L[0]);
      double ly = L[1] * std::sqrt(rx);
      if (almost_integer(ly)) { retry = true; break; }
      for (int ky = 0; ky <= (int)ly; ky++) {
        double ry = rx - (ky / L[1]) * (ky / L[1]);
        double lz = L[2] * std::sqrt(ry);
        if (almost_integer(lz)) { retry = true; break; }
        for (int kz = (kx + ky) ? 0 : 1; kz <= (int)lz; kz++)
          count += 1L << ((kx > 0) + (ky > 0) + (kz > 0) - 1);
      }
    }
    if (!retry) { *nk = (int)count; return 0; }
    *kappa += *kappa * dK;
    dK *= 1.4;
  }
  return mdx::fail(MDX_E_NOCONV, "mdx_ewald_kvectors: cannot resolve rounding of the k-limits");
}

// lcsetup(): number of linked cells per direction from the target of
// 1+ln(No.s) sites per cell (cook/linklist.c:304-337, serial build)
extern "C" int mdx_lcsetup(int ncell[3], int ns, double percell, const double L[3]) {
  if (!ncell || ns <= 0) return mdx::fail(MDX_E_ARG, "mdx_lcsetup: bad arguments");
  if (percell <= 0) percell = 1 + std::log((double)ns);
  int nundef = 0, prdc = 1;
  double prdL = 1;
  for (int i = 0; i < 3; i++) {
    if (ncell[i] == 0) { nundef++; prdL *= L[i]; }
    else prdc *= ncell[i];
  }
  for (int i = 0; i < 3; i++) {
    if (ncell[i] != 0) continue;
    double x = std::pow(ns / (percell * prdc * prdL), 1.0 / nundef);
    ncell[i] = logint(x * L[i]);
    if (ncell[i] < 2) ncell[i] = 2;
    prdL /= L[i];
    prdc *= ncell[i];
    nundef--;
  }
  return 0;
}

// cutcorr() (sim/siminit.c:665-722): the homogeneous LJ cutoff correction
// summed over all intermolecular site pairs, En.corr = sum_{molecule pairs
// a<b} sum_{i in a, j in b} corr(type_i, type_j).  Written over site-type
// counts: 1/2 (N.C.N - sum_a n_a.C.n_a) with N the counts of the system and
// n_a those of molecule a; the same sum as the reference's species loop
// with N(N-1)/2 pairs of equal species.  corr: [ntypes*ntypes] from
// mdx_setss.  Energy correction Ecorr/V, pressure correction Ecorr/V^2.
extern "C" int mdx_cutcorr(const mdx_system_t *s, const double *corr, double *Ecorr) {
  if (!s || !corr || !Ecorr || !s->mol_first || !s->site_type || s->ntypes <= 0)
    return mdx::fail(MDX_E_ARG, "mdx_cutcorr: NULL argument");
  const int nt = s->ntypes;
  std::vector<double> N(nt, 0.0), n(nt, 0.0);
  for (int i = 0; i < s->ns; i++) {
    const int t = s->site_type[i];
    if (t < 0 || t >= nt) return mdx::fail(MDX_E_ARG, "mdx_cutcorr: site type out of range");
    N[t] += 1;
  }
  auto quad = [&](const std::vector<double> &v) {
    double q = 0;
    for (int a = 0; a < nt; a++)
      for (int b = 0; b < nt; b++) q += v[a] * corr[a * nt + b] * v[b];
    return q;
  };
  double intra = 0;
  for (int m = 0; m < s->nmol; m++) {
    std::fill(n.begin(), n.end(), 0.0);
    for (int i = s->mol_first[m]; i < s->mol_first[m + 1]; i++) n[s->site_type[i]] += 1;
    intra += quad(n);
  }
  *Ecorr = 0.5 * (quad(N) - intra);
  return 0;
}

// measurePconstraints() (sim/constrd.c:115-265, PRESSURETENSOR=7, no ECC,
// no slab): the virial sum and the pressure tensors of one measured step.
extern "C" int mdx_pressure(const mdx_step_out_t *so, const double L[3], double Pkinq,
                            double Ecorr, int polar, mdx_pressure_t *p) {
  if (!so || !L || !p) return mdx::fail(MDX_E_ARG, "mdx_pressure: NULL argument");
  const double V = L[0] * L[1] * L[2];
  if (!(V > 0)) return mdx::fail(MDX_E_ARG, "mdx_pressure: box volume must be > 0");
  std::memset(p, 0, sizeof(*p));
  // En.vir += En.virc + En.el (- 2 En.self with Drudes)
  p->vir = so->frc.En_vir + so->virc + so->frc.En_el - (polar ? 2 * so->En_self : 0.0);
  for (int i = 0; i < 6; i++) {
    p->Pkin[i] = so->Pkin[i] * Pkinq / V;
    p->Pvir[i] = (so->frc.Pvir[i] + so->PTvirc[i]) / V;
    p->Ptens[i] = p->Pvir[i] + p->Pkin[i];
  }
  const double cP = Ecorr / (V * V);
  p->Ptr = (p->Ptens[0] + p->Ptens[1] + p->Ptens[2]) / 3;
  p->Ptr_c = p->Ptr + cP;
  // En.kin is the doubled kinetic energy at this point of the reference
  p->Pevir = (2 * so->En_kin * Pkinq + p->vir) / 3 / V;
  p->Pevir_c = p->Pevir + cP;
  return 0;
}

namespace mdx {

// Ediag, sim/ewald.c:477-492: the omitted k-space terms with kappa^2 <
// k^2 < kappa^2 + alpha^2 over the full k lattice, each exp(-pi^2 k^2 /
// alpha^2)/k^2, times sumq/(2 pi V); cook truncates K = (int)(kappax L)
double ewald_ediag(const double L[3], double kappa, double alpha, double sumq) {
  const double pi = 3.14159265358979323846;
  const double qkx = kappa * kappa + alpha * alpha, kx_ = std::sqrt(qkx);
  const double q = -(pi / alpha) * (pi / alpha), kq = kappa * kappa;
  const int K[3] = {(int)(kx_ * L[0]), (int)(kx_ * L[1]), (int)(kx_ * L[2])};
  double E = 0;
  for (int kx = -K[0]; kx <= K[0]; kx++)
    for (int ky = -K[1]; ky <= K[1]; ky++)
      for (int kz = -K[2]; kz <= K[2]; kz++) {
        const double a = kx / L[0], b = ky / L[1], c = kz / L[2];
        const double kk = a * a + b * b + c * c;
        if (kk > kq && kk < qkx) E += std::exp(kk * q) / kk;
      }
  return E * sumq / (2 * pi * L[0] * L[1] * L[2]);
}

void ewald_dipole_coefs(int el_corr, int el_bg, double epsinf, double qnet, double V,
                        double cm[3], double ce[3], double cb[3]) {
  const double pi = 3.14159265358979323846;
  for (int d = 0; d < 3; d++) cm[d] = ce[d] = cb[d] = 0;
  if (el_corr) {
    // Yeh-Berkowitz in the directions of el.corr (sim/ewald.c:869-882);
    // +8: energy and pressure only; with el.bg the background field
    // 4 pi Q r_d / V in the same directions (sim/ewald.c:901-931)
    const double c = 4 * pi / V;
    for (int d = 0; d < 3; d++)
      if (el_corr & 1 << d) {
        ce[d] = c;
        if (!(el_corr & 8)) cm[d] = c;
        if (el_bg) cb[d] = 4 * pi * qnet / V;
      }
  } else {
    // dielectric boundary epsinf (sim/ewald.c:884-887)
    const double c = 4 * pi / (2 * epsinf + 1) / V;
    for (int d = 0; d < 3; d++) cm[d] = ce[d] = c;
  }
}

int ewald_check(int el_corr, int el_bg, double epsinf, double qnet, double sumq2,
                const char *who) {
  if (el_corr < 0 || el_corr > 15)
    return fail(MDX_E_ARG, std::string(who) + ": el_corr must be in 0..15");
  if (el_bg < 0 || el_bg > 2) return fail(MDX_E_ARG, std::string(who) + ": el_bg must be 0, 1 or 2");
  if (el_corr && std::fabs(epsinf) < 1e15)
    return fail(MDX_E_ARG, std::string(who) +
                               ": the Yeh-Berkowitz correction (el_corr) needs epsinf = infinity");
  // cook's neutrality test Q/sqrt(sumq) > el.epsq (sim/siminit.c:152-160,
  // sim/ewaldprt.c:16): el.epsq = 0 there, a relative charge below 1e-10
  // counts as rounding here
  if (qnet * qnet > 1e-20 * sumq2) {
    if (!el_bg)
      return fail(MDX_E_ARG, std::string(who) + ": the system is charged and el_bg = 0");
    if (std::fabs(epsinf) < 1e16)
      return fail(MDX_E_ARG, std::string(who) +
                                 ": a charged system needs a tin-foil epsinf (|epsinf| >= 1e16)");
  }
  return 0;
}

}  // namespace mdx
