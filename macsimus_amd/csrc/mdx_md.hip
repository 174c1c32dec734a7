// Device-resident MD step: leap-frog Verlet + SHAKE per molecule, kinetic
// energy/tensor, constraint virial and the Berendsen thermostat, following
// Shake() of sim/constrd.c:622-1081 for SHAKE=1 (simplified formula,
// sim/constrd.c:805-839), VERLET=3 (sim/shakev3.c), sim/verlet.c (no
// thermostat branch), sim/virc.c and sim/thermo.c:156-160.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "mdx_internal.h"

namespace mdx {
int forces_device(Ctx &c, int measure);
void fill_out(const Ctx &c, const double *acc, int measure, mdx_out_t *out);
}  // namespace mdx

using namespace mdx;

namespace {

constexpr int kBlock = 256;
constexpr int kAspcMax = 8;          // longest ASPC predictor (cook warns above 8)

__global__ void k_fill(long n, double *__restrict__ x, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}
enum { MD_EKIN = 0, MD_PKIN = 1, MD_VIRC = 7, MD_PTV = 8, MD_KTR = 14, MD_ITMAX = 15, MD_FAIL = 16,
       MD_N = 17 };

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// the time-reversible velocity predictor of one step (sim/trvpscale.c):
// coefficients b[0..kk] for the history known so far, the length knew
// after this step
struct Trvp {
  int kk, knew;
  double b[10];
};

Trvp trvp_coef(int kk, int kmax) {
  Trvp t{};
  t.kk = kk;
  t.knew = std::min(kk + 1, kmax);
  t.b[0] = (double)(2 * kk + 1) / (kk + 1);
  for (int j = 0; j < kk; j++) t.b[j + 1] = -t.b[j] * (double)(kk - j) / (kk + j + 2);
  return t;
}

// one thread per molecule: p = h^2 f/m + r + r1, SHAKE sweeps until the
// largest relative violation is <= eps, then r1 = p - r, r = p.
// The velocity predictor (pred != PRED_NONE, sim/verlet.c:22-36): p -=
// Vfactor vpred with the predicted h v of each site from b and the history
// (level 0 = the current r1, level j = hist[j-1], stride 1+3ns: a logs
// slot then the sites); Vfactor is h dlogs/dt predicted from the device
// Nose variables (PRED_NOSE: nose[1] and the logs slots) or given by the
// host (PRED_HOST, the NPT Vfactor of sim/trvpscale.c:62).  Each thread
// then shifts its sites' history (the logs slots are shifted by k_nose
// after all threads have read them).  bscale multiplies p_i - p_j in the
// constraint (the predicted box scaling of NPT, sim/constrd.c:820; 1 else).
enum { PRED_NONE = 0, PRED_NOSE = 1, PRED_HOST = 2 };
__global__ __launch_bounds__(kBlock) void k_shake(
    int nmol, const int *__restrict__ molfirst, const int *__restrict__ cons_first,
    const int *__restrict__ cons_ij, const double *__restrict__ bondq,
    const double *__restrict__ omega, const double *__restrict__ mass, double h, double eps,
    int maxit, int measure, double *__restrict__ r, double *__restrict__ r1,
    const double *__restrict__ f, double *__restrict__ p, double *__restrict__ part, int pred,
    const double *__restrict__ nose, double vfac_host, double bscale, double *__restrict__ hist,
    long stride, Trvp tv, const int *__restrict__ msp, unsigned long long *__restrict__ nit) {
  const int m = blockIdx.x * kBlock + threadIdx.x;
  double ek = 0, pk[6] = {0, 0, 0, 0, 0, 0}, virc = 0, ptv[6] = {0, 0, 0, 0, 0, 0};
  double its = 0, bad = 0, ktr = 0;
  unsigned long long nsw = 0;          // sweeps of a molecule with constraints
  int sp = -1;                         // its species
  if (m < nmol) {
    const double hh = h * h;
    const int b = molfirst[m], e = molfirst[m + 1];
    if (pred != PRED_NONE) {
      double vfac = vfac_host;
      if (pred == PRED_NOSE) {
        vfac = tv.b[0] * nose[1];
        for (int j = 1; j <= tv.kk; j++) vfac += tv.b[j] * hist[(j - 1) * stride];
      }
      for (int i = b; i < e; i++) {
        const double im = hh / mass[i];
        for (int k = 0; k < 3; k++) {
          const long x = 3L * i + k;
          double v = tv.b[0] * r1[x];
          for (int j = 1; j <= tv.kk; j++) v += tv.b[j] * hist[(j - 1) * stride + 1 + x];
          p[x] = im * f[x] - vfac * v + r[x] + r1[x];
          for (int j = tv.knew; j >= 2; j--) hist[(j - 1) * stride + 1 + x] = hist[(j - 2) * stride + 1 + x];
          if (tv.knew >= 1) hist[1 + x] = r1[x];
        }
      }
    } else {
      for (int i = b; i < e; i++) {
        const double im = hh / mass[i];
        for (int k = 0; k < 3; k++) p[3 * i + k] = im * f[3 * i + k] + r[3 * i + k] + r1[3 * i + k];
      }
    }
    const int c0 = cons_first[m], c1 = cons_first[m + 1], nc = c1 - c0;
    const double om = omega[m];
    const int mit = maxit > 0 ? maxit : (int)(-16 * log(eps * 0.1)) * (nc + 4);
    int it = 0;
    if (nc) {
      double maxsq;
      do {
        it++;
        maxsq = 0;
        for (int al = c0; al < c1; al++) {
          const int a = cons_ij[2 * al], c = cons_ij[2 * al + 1];
          const double bq = bondq[al], mi = 1 / mass[a], mj = 1 / mass[c];
          double pij[3], rij[3];
          for (int k = 0; k < 3; k++) {
            pij[k] = (p[3 * a + k] - p[3 * c + k]) * bscale;
            rij[k] = r[3 * a + k] - r[3 * c + k];
          }
          double sq = bq - (pij[0] * pij[0] + pij[1] * pij[1] + pij[2] * pij[2]);
          maxsq = fmax(maxsq, fabs(sq / bq));
          sq *= om / (mi + mj);
          if (measure) {
            const double w = sq / bq;
            ptv[0] += w * rij[0] * rij[0]; ptv[1] += w * rij[1] * rij[1]; ptv[2] += w * rij[2] * rij[2];
            ptv[3] += w * rij[1] * rij[2]; ptv[4] += w * rij[2] * rij[0]; ptv[5] += w * rij[0] * rij[1];
          }
          virc += sq;
          sq /= bq;
          for (int k = 0; k < 3; k++) {
            const double d = rij[k] * sq;
            p[3 * a + k] += mi * d;
            p[3 * c + k] -= mj * d;
          }
        }
      } while (maxsq > eps && it < mit);
      if (maxsq > eps) bad = 1;
    }
    its = it;
    nsw = nc ? it : 0;
    sp = msp ? msp[m] : 0;
    double vm1[3] = {0, 0, 0}, vm2[3] = {0, 0, 0}, mm = 0;
    for (int i = b; i < e; i++) {
      double V1[3], V2[3];
      for (int k = 0; k < 3; k++) { V1[k] = r1[3 * i + k]; V2[k] = p[3 * i + k] - r[3 * i + k]; }
      const double mi = mass[i];
      for (int k = 0; k < 3; k++) { vm1[k] += mi * V1[k]; vm2[k] += mi * V2[k]; }
      mm += mi;
      ek += (V1[0] * V1[0] + V1[1] * V1[1] + V1[2] * V1[2] + V2[0] * V2[0] + V2[1] * V2[1] + V2[2] * V2[2]) * mi;
      pk[0] += (V1[0] * V1[0] + V2[0] * V2[0]) * mi;
      pk[1] += (V1[1] * V1[1] + V2[1] * V2[1]) * mi;
      pk[2] += (V1[2] * V1[2] + V2[2] * V2[2]) * mi;
      pk[3] += (V1[1] * V1[2] + V2[1] * V2[2]) * mi;
      pk[4] += (V1[2] * V1[0] + V2[2] * V2[0]) * mi;
      pk[5] += (V1[0] * V1[1] + V2[0] * V2[1]) * mi;
      for (int k = 0; k < 3; k++) { r1[3 * i + k] = V2[k]; r[3 * i + k] = p[3 * i + k]; }
    }
    // translational kinetic energy of the molecule (sim/shakev3.c:35)
    ktr = (vm1[0] * vm1[0] + vm1[1] * vm1[1] + vm1[2] * vm1[2] +
           vm2[0] * vm2[0] + vm2[1] * vm2[1] + vm2[2] * vm2[2]) / mm;
  }
  // sweep counts per species (cook's constrit nit, sim/constrd.c:930):
  // one atomic per wave where the wave's molecules share a species (the
  // molecules of a species are consecutive), else one per molecule;
  // integer sums, so the order does not matter
  {
    const int s0 = __shfl(sp, 0, 64);
    if (__all(sp == s0 || sp < 0)) {
      unsigned long long t = nsw;
      for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
      if ((threadIdx.x & 63) == 0 && t && s0 >= 0) atomicAdd(nit + s0, t);
    } else if (sp >= 0 && nsw) {
      atomicAdd(nit + sp, nsw);
    }
  }
  __shared__ double red[kBlock / 64][MD_N];
  const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double v[MD_N] = {ek, pk[0], pk[1], pk[2], pk[3], pk[4], pk[5], virc,
                    ptv[0], ptv[1], ptv[2], ptv[3], ptv[4], ptv[5], ktr, its, bad};
  for (int k = 0; k < MD_N; k++) {
    const double s = k >= MD_ITMAX ? wave_max(v[k]) : wave_sum(v[k]);
    if (ln == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < MD_N) {
    double s = threadIdx.x >= MD_ITMAX ? 0 : 0;
    for (int w = 0; w < kBlock / 64; w++)
      s = threadIdx.x >= MD_ITMAX ? fmax(s, red[w][threadIdx.x]) : s + red[w][threadIdx.x];
    part[(size_t)blockIdx.x * MD_N + threadIdx.x] = s;
  }
}

// flag[1] = SHAKE failed and flag[2] = max SHAKE iterations, both sticky
// over the steps of one mdx_md_steps call
__global__ void k_md_reduce(const double *part, int nblocks, double *acc, double *flag) {
  __shared__ double red[kBlock];
  for (int a = 0; a < MD_N; a++) {
    const bool mx = a >= MD_ITMAX;
    double s = 0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock)
      s = mx ? fmax(s, part[(size_t)b * MD_N + a]) : s + part[(size_t)b * MD_N + a];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o)
        red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + o])
                              : red[threadIdx.x] + red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      acc[a] = red[0];
      if (a == MD_FAIL) flag[1] = fmax(flag[1], red[0]);
      if (a == MD_ITMAX) flag[2] = fmax(flag[2], red[0]);
    }
    __syncthreads();
  }
}

// Berendsen: r1 *= exp(blogT(T_kin, T) * -h/tau) with blogT = ln(E/T)/2
// clamped to [-1,1] (sim/constrd.c:98-107); T_kin = 2 E_kin / No.f with
// acc[MD_EKIN] = sum m (V1^2+V2^2) = 4 h^2 E_kin (leap-frog estimate)
__global__ void k_berendsen(int n3, double *r1, const double *acc, double h, double T,
                            double tau, double nf, double *lambda_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double ekin2 = acc[MD_EKIN] * 0.5 / (h * h);  // En.kin (doubled) after hscale
  const double Tk = ekin2 / nf;
  const double e2 = 7.38905609893065022723;
  const double b = Tk > T * e2 ? 1 : (Tk * e2 < T ? -1 : log(Tk / T) / 2);
  const double lam = exp(b * -h / tau);
  if (i < n3) r1[i] *= lam;
  if (i == 0 && lambda_out) *lambda_out = lam;
}

// the isotropic NPT rescaling of the sites (rescalecfg(cfg[0], RESCALE_XYZ
// | RESCALE_L, 0, scale), sim/norm.c:989-1000): r(t+h) *= scale; r1 stays
__global__ void k_rescale(int n3, double *r, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) r[i] *= scale;
}

// Nose-Hoover, after SHAKE (sim/thermo.c:180-200): shifts the logs slots
// of the predictor history, then the leap-frog of logs with the doubled
// kinetic energy T_kin = 2 E_kin / No.f and the VERLET=3 extended energy
// En.ext = No.f T (logs(t) + (tau/h)^2 (dlogs_old^2 + dlogs_new^2)/4)
__global__ void k_nose(const double *acc, double *nose, double *hist, long stride, int knew,
                       double h, double T, double tau, double nf) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double hh = h * h;
  const double Tk = acc[MD_EKIN] * 0.5 / hh / nf;
  const double ld = nose[1], logs0 = nose[0];
  for (int j = knew; j >= 2; j--) hist[(j - 1) * stride] = hist[(j - 2) * stride];
  if (knew >= 1) hist[0] = ld;
  nose[1] = ld + hh * (Tk / T - 1) / (tau * tau);
  nose[0] = logs0 + nose[1];
  nose[3] = (ld + nose[1]) / (2 * h);
  nose[2] = (logs0 + (tau / h) * (tau / h) * (ld * ld + nose[1] * nose[1]) / 4) * nf * T;
}

// the decoupled friction thermostats after SHAKE (sim/thermo.c:107-135),
// one thread per molecule: ze acts on the molecule's translation (kind
// 11), z on the intramolecular motion (12), both with their own
// temperatures from T_tr_in (13); r1 -= z r1 + (ze - z) v_mol with the
// linear factors blogT(T_kin/T) h/tau of the doubled kinetic energies
__device__ inline double blogT(double E, double T) {
  const double e2 = 7.38905609893065022723;
  return E > T * e2 ? 1 : (E * e2 < T ? -1 : log(E / T) / 2);
}

__global__ __launch_bounds__(kBlock) void k_friction(int nmol, const int *__restrict__ molfirst,
                                                     const double *__restrict__ mass,
                                                     double *__restrict__ r1, const double *acc,
                                                     double h, double T, double tau, double nf,
                                                     double nf_tr, double T_tr_in, int kind) {
  const int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= nmol) return;
  const double hs = 0.5 / (h * h);
  const double kin = acc[MD_EKIN] * hs, ktr = acc[MD_KTR] * hs, f_in = nf - nf_tr;
  double z = 0, ze = 0;
  if (kind == 11) ze = blogT(ktr / nf_tr, T) * (h / tau);
  if (kind == 12) z = blogT((kin - ktr) / f_in, T) * (h / tau);
  if (kind == 13) {
    const double T_tr = (f_in + nf_tr) * T / (T_tr_in * f_in + nf_tr);
    const double T_in = (f_in + nf_tr) * T / (f_in + nf_tr / T_tr_in);
    z = blogT((kin - ktr) / f_in, T_in) * (h / tau);
    ze = blogT(ktr / nf_tr, T_tr) * (h / tau);
  }
  ze -= z;
  const int b = molfirst[m], e = molfirst[m + 1];
  double vm[3] = {0, 0, 0}, mm = 0;
  for (int i = b; i < e; i++) {
    for (int k = 0; k < 3; k++) vm[k] += mass[i] * r1[3 * i + k];
    mm += mass[i];
  }
  for (int k = 0; k < 3; k++) vm[k] /= mm;
  for (int i = b; i < e; i++)
    for (int k = 0; k < 3; k++) r1[3 * i + k] -= z * r1[3 * i + k] + ze * vm[k];
}

// ------------------------------------------------------------------ Drude
enum { SF_EEL = 0, SF_PV = 1, SF_SELF = 7, SF_ERR = 8, SF_BAD = 9, SF_N = 10 };

// exacterud_sqrt_1(): -erf(alpha r)/r + 2 alpha/sqrt(pi) and its force/r,
// Ewald correction of the centre-Drude pair of one site (sim/erexact.c:55-82)
__device__ inline double exact_erud1(double rr, double alpha, double &erd) {
  const double alphaq = alpha * alpha, A = alpha * (-2 / 1.772453850905516), B = A * alphaq * 2;
  const double err = alphaq * rr;
  if (err < 0.01) {
    double su = 0, sd = 0, q = 1;
    for (int n = 1;;) {
      const double t = q / (2 * n + 1);
      sd += t * n;
      su += t;
      if (fabs(q) < 1e-17) break;
      q *= -err / ++n;
    }
    erd = B * sd;
    return A * (-su * err);
  }
  const double r = sqrt(rr), y = -erf(alpha * r) / r;
  erd = (y - A * exp(-err)) / rr;
  return y - A;
}

// one selffield() iteration per Drude site (sim/scf.c:120-240), FP64:
// r, f with 2ns rows (centres, Drude offsets / forces)
__global__ __launch_bounds__(kBlock) void k_selffield(
    int ns, double *__restrict__ r, double *__restrict__ f, const double *__restrict__ q,
    const double *__restrict__ qpol, const double *__restrict__ alpha_qq, double alpha,
    double omega, double minqq2, double *__restrict__ part) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[SF_N] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < ns && qpol[i] != 0) {
    double *rp = r + 3 * (ns + i), *fp = f + 3 * (ns + i), *fc = f + 3 * i;
    const double d[3] = {rp[0], rp[1], rp[2]};
    const double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (rr >= minqq2) v[SF_BAD] = 1;
    double erd;
    double ff = q[i] * qpol[i];
    v[SF_EEL] = ff * exact_erud1(rr, alpha, erd);
    ff *= erd;
    double p[3] = {fp[0], fp[1], fp[2]}, c[3] = {fc[0], fc[1], fc[2]};
    // MADDFORCES(fpol, fc, ff, d) with its tensor order
    double fdr;
    p[0] += fdr = ff * d[0]; c[0] -= fdr; v[SF_PV + 0] += d[0] * fdr; v[SF_PV + 5] += d[1] * fdr;
    p[1] += fdr = ff * d[1]; c[1] -= fdr; v[SF_PV + 1] += d[1] * fdr; v[SF_PV + 3] += d[2] * fdr;
    p[2] += fdr = ff * d[2]; c[2] -= fdr; v[SF_PV + 2] += d[2] * fdr; v[SF_PV + 4] += d[0] * fdr;
    double nr[3], dd2 = 0;
    for (int k = 0; k < 3; k++) {
      nr[k] = (1 - omega) * d[k] + omega * alpha_qq[i] * p[k];
      dd2 += (d[k] - nr[k]) * (d[k] - nr[k]);
      c[k] += p[k];
      rp[k] = nr[k];
      fp[k] = p[k];
      fc[k] = c[k];
    }
    v[SF_SELF] = nr[0] * p[0] + nr[1] * p[1] + nr[2] * p[2];
    v[SF_PV + 0] -= nr[0] * p[0];
    v[SF_PV + 1] -= nr[1] * p[1];
    v[SF_PV + 2] -= nr[2] * p[2];
    v[SF_PV + 3] -= nr[1] * p[2];
    v[SF_PV + 4] -= nr[2] * p[0];
    v[SF_PV + 5] -= nr[0] * p[1];
    v[SF_ERR] = dd2 / (omega * omega) * qpol[i] * qpol[i];
    if (nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2] >= minqq2) v[SF_BAD] = 1;
  }
  __shared__ double red[kBlock / 64][SF_N];
  const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < SF_N; k++) {
    const double s = k >= SF_ERR ? wave_max(v[k]) : wave_sum(v[k]);
    if (ln == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < SF_N) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++)
      s = threadIdx.x >= SF_ERR ? fmax(s, red[w][threadIdx.x]) : s + red[w][threadIdx.x];
    part[(size_t)blockIdx.x * SF_N + threadIdx.x] = s;
  }
}

// flag[3] = a Drude left el.minqq (sticky over the steps of one call)
__global__ void k_sf_reduce(const double *part, int nblocks, double *acc, double *flag) {
  __shared__ double red[kBlock];
  for (int a = 0; a < SF_N; a++) {
    const bool mx = a >= SF_ERR;
    double s = 0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock)
      s = mx ? fmax(s, part[(size_t)b * SF_N + a]) : s + part[(size_t)b * SF_N + a];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o)
        red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + o])
                              : red[threadIdx.x] + red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      acc[a] = red[0];
      if (a == SF_BAD) flag[3] = fmax(flag[3], red[0]);
    }
    __syncthreads();
  }
}

// ASPC predictor + history shift of sim/pred2.c:54-64: aa[0] = Drude rows
// of r (rpol(t)), aa[1] = Drude rows of r1, aa[2..k+1] = hist; the
// prediction sum_{i<kk+2} B_i aa[i] replaces aa[0] after aa[i+1] <- aa[i]
struct AspcB {
  double b[kAspcMax + 2];
};

__global__ void k_aspc(int n3, double *__restrict__ a0, double *__restrict__ a1,
                       double *__restrict__ hist, int k, int kk, AspcB B) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n3) return;
  double aa[kAspcMax + 2];
  aa[0] = a0[j];
  aa[1] = a1[j];
#pragma unroll
  for (int m = 2; m < kAspcMax + 2; m++) aa[m] = m <= k + 1 ? hist[(size_t)(m - 2) * n3 + j] : 0.0;
  double pred = 0;
#pragma unroll
  for (int m = 0; m < kAspcMax + 2; m++) pred += m < kk + 2 ? B.b[m] * aa[m] : 0.0;
  // shift: aa[m+1] <- aa[m] for m = k..0 (aa[k+1] drops out)
#pragma unroll
  for (int m = 2; m < kAspcMax + 2; m++)
    if (m <= k + 1) hist[(size_t)(m - 2) * n3 + j] = aa[m - 1];
  a1[j] = aa[0];
  a0[j] = pred;
}

// B_1..B_{kk+2} of the ASPC predictor of length kk (sim/pred2.c:94-98;
// kk = 0 gives the 2 a0 - a1 start of sim/pred2.c:48)
AspcB aspc_coef(int kk) {
  AspcB B{};
  double o = (double)(4 * kk + 6) / (double)(kk + 3);
  for (int i = 0; i < kk + 2; i++) {
    B.b[i] = o * (i + 1);
    o *= (double)(i - kk - 1) / (double)(kk + i + 4);
  }
  return B;
}

void launch_selffield(Ctx &c, double omega) {
  const int nb = nblk(c.ns);
  k_selffield<<<nb, kBlock, 0, c.stream>>>(c.ns, c.r.p, c.f.p, c.charge.p, c.chargepol.p,
                                          c.alpha_qq.p, c.alphar, omega, 1.0, c.sfpart.p);
  k_sf_reduce<<<1, kBlock, 0, c.stream>>>(c.sfpart.p, nb, c.sfacc.p, c.mdflag.p);
}

int set_aspc(Ctx &c, int k, int kk, const double *hist) {
  if (k < 0 || k > kAspcMax || kk < 0 || kk > k)
    return fail(MDX_E_ARG, "ASPC: need 0 <= kk <= k <= 8");
  const size_t n = (size_t)std::max(k, 1) * 3 * c.ns;
  int rc;
  if (c.aspc_k != k || !c.aspc_hist.p) {
    c.aspc_hist.release();
    if ((rc = c.aspc_hist.alloc(n))) return rc;
    c.gen++;
  }
  if (hist)
    MDX_HIP(hipMemcpyAsync(c.aspc_hist.p, hist, sizeof(double) * 3 * c.ns * k,
                           hipMemcpyHostToDevice, c.stream));
  else
    k_fill<<<nblk((long)n), kBlock, 0, c.stream>>>((long)n, c.aspc_hist.p, 0.0);
  MDX_HIP(hipGetLastError());
  MDX_HIP(hipStreamSynchronize(c.stream));
  c.aspc_k = k;
  c.aspc_kk = kk;
  return 0;
}

// the predictor of the Nose-Hoover step: length k (0..9), reached kk,
// history [k][1+3ns] (NULL: zeros)
int set_trvp(Ctx &c, int k, int kk, const double *hist) {
  if (k < 0 || k > 9 || kk < 0 || kk > k) return fail(MDX_E_ARG, "Nose: need 0 <= kk <= k <= 9");
  const size_t n = (size_t)std::max(k, 1) * (1 + 3 * (size_t)c.ns);
  int rc;
  if (c.trvp_k != k || !c.trvp_hist.p) {
    c.trvp_hist.release();
    if ((rc = c.trvp_hist.alloc(n))) return rc;
    c.gen++;
  }
  if (hist && k > 0)
    MDX_HIP(hipMemcpyAsync(c.trvp_hist.p, hist, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
  else
    k_fill<<<nblk((long)n), kBlock, 0, c.stream>>>((long)n, c.trvp_hist.p, 0.0);
  MDX_HIP(hipGetLastError());
  MDX_HIP(hipStreamSynchronize(c.stream));
  c.trvp_k = k;
  c.trvp_kk = kk;
  return 0;
}

}  // namespace

int mdx::md_setup(Ctx &c) {
  int rc;
  if ((rc = c.mdflag.alloc(4)) || (rc = c.shk_nit.alloc(64))) return rc;
  if (c.mass.p && ((rc = c.mdpart.alloc((size_t)nblk(c.nmol) * MD_N)) || (rc = c.mdacc.alloc(MD_N)) ||
                    (rc = c.nose.alloc(4))))
    return rc;
  if (c.mass.p) {
    k_fill<<<1, 64, 0, c.stream>>>(4, c.nose.p, 0.0);
    MDX_HIP(hipGetLastError());
  }
  if (c.polar && ((rc = c.sfpart.alloc((size_t)nblk(c.ns) * SF_N)) || (rc = c.sfacc.alloc(SF_N))))
    return rc;
  return 0;
}

namespace {
// flags: lambda = 1, the sticky ones 0 (a kernel, so that the step can be captured)
__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

void reset_flags(Ctx &c) {
  k_fill<<<1, 64, 0, c.stream>>>(4, c.mdflag.p, 0.0);
  k_fill<<<1, 64, 0, c.stream>>>(1, c.mdflag.p, 1.0);
  k_zero_u64<<<1, 64, 0, c.stream>>>(c.shk_nit.p, c.shk_nsp);
}
}  // namespace

extern "C" int mdx_selffield(mdx_ctx_t *h, double *r, double *f, mdx_out_t *out, double omega,
                             double *Eself, double *maxerr) {
  if (!h || !r || !f || !out) return fail(MDX_E_ARG, "mdx_selffield: NULL argument");
  Ctx &c = h->c;
  if (!c.polar) return fail(MDX_E_ARG, "mdx_selffield: system has no Drude charges");
  if (!(omega > 0 && omega <= 1)) return fail(MDX_E_ARG, "mdx_selffield: omega in (0,1]");
  MDX_HIP(hipSetDevice(c.device));
  const size_t bytes = sizeof(double) * 6 * c.ns;
  double sf[SF_N], flag[4];
  MDX_HIP(hipMemcpyAsync(c.r.p, r, bytes, hipMemcpyHostToDevice, c.stream));
  MDX_HIP(hipMemcpyAsync(c.f.p, f, bytes, hipMemcpyHostToDevice, c.stream));
  reset_flags(c);
  launch_selffield(c, omega);
  MDX_HIP(hipGetLastError());
  MDX_HIP(hipMemcpyAsync(sf, c.sfacc.p, sizeof(sf), hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipMemcpyAsync(flag, c.mdflag.p, sizeof(flag), hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipMemcpyAsync(r, c.r.p, bytes, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipMemcpyAsync(f, c.f.p, bytes, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipStreamSynchronize(c.stream));
  c.state_on_device = false;           // r now holds the caller's configuration
  if (flag[3] > 0) return fail(MDX_E_RANGE, "selffield: Drude farther than el.minqq=1 A from its centre");
  out->En_el += sf[SF_EEL];
  for (int k = 0; k < 6; k++) out->Pvir[k] += sf[SF_PV + k];
  if (Eself) *Eself = sf[SF_SELF] / 2;
  if (maxerr) *maxerr = sqrt(sf[SF_ERR]);
  return 0;
}

extern "C" int mdx_upload_state(mdx_ctx_t *h, const double *r, const double *r1) {
  if (!h || !r || !r1) return fail(MDX_E_ARG, "mdx_upload_state: NULL argument");
  Ctx &c = h->c;
  if (!c.mass.p) return fail(MDX_E_ARG, "mdx_upload_state: system has no masses");
  MDX_HIP(hipSetDevice(c.device));
  const size_t bytes = sizeof(double) * 3 * (c.polar ? 2 * c.ns : c.ns);
  MDX_HIP(hipMemcpyAsync(c.r.p, r, bytes, hipMemcpyHostToDevice, c.stream));
  MDX_HIP(hipMemcpyAsync(c.r1.p, r1, bytes, hipMemcpyHostToDevice, c.stream));
  MDX_HIP(hipStreamSynchronize(c.stream));
  c.state_on_device = true;
  return 0;
}

extern "C" int mdx_download_state(mdx_ctx_t *h, double *r, double *r1) {
  if (!h || !r || !r1) return fail(MDX_E_ARG, "mdx_download_state: NULL argument");
  Ctx &c = h->c;
  if (!c.state_on_device) return fail(MDX_E_ARG, "mdx_download_state: no state uploaded");
  MDX_HIP(hipSetDevice(c.device));
  const size_t bytes = sizeof(double) * 3 * (c.polar ? 2 * c.ns : c.ns);
  MDX_HIP(hipMemcpyAsync(r, c.r.p, bytes, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipMemcpyAsync(r1, c.r1.p, bytes, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

extern "C" int mdx_set_aspc(mdx_ctx_t *h, int k, int kk, const double *hist) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_aspc: NULL handle");
  Ctx &c = h->c;
  if (!c.polar) return fail(MDX_E_ARG, "mdx_set_aspc: system has no Drude charges");
  MDX_HIP(hipSetDevice(c.device));
  return set_aspc(c, k, kk, hist);
}

extern "C" int mdx_get_aspc(mdx_ctx_t *h, int *k, int *kk, double *hist) {
  if (!h || !k || !kk) return fail(MDX_E_ARG, "mdx_get_aspc: NULL argument");
  Ctx &c = h->c;
  if (!c.polar) return fail(MDX_E_ARG, "mdx_get_aspc: system has no Drude charges");
  *k = c.aspc_k;
  *kk = c.aspc_kk;
  if (hist && c.aspc_k > 0) {
    MDX_HIP(hipSetDevice(c.device));
    MDX_HIP(hipMemcpyAsync(hist, c.aspc_hist.p, sizeof(double) * 3 * c.ns * c.aspc_k,
                           hipMemcpyDeviceToHost, c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
  }
  return 0;
}

extern "C" int mdx_set_nose(mdx_ctx_t *h, double logs, double logs1, int k, int kk,
                            const double *hist) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_nose: NULL handle");
  Ctx &c = h->c;
  if (!c.mass.p) return fail(MDX_E_ARG, "mdx_set_nose: system has no masses");
  MDX_HIP(hipSetDevice(c.device));
  int rc;
  if ((rc = set_trvp(c, k, kk, hist))) return rc;
  const double v[4] = {logs, logs1, 0, 0};
  MDX_HIP(hipMemcpyAsync(c.nose.p, v, sizeof(v), hipMemcpyHostToDevice, c.stream));
  MDX_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

extern "C" int mdx_get_nose(mdx_ctx_t *h, double *logs, double *logs1, int *k, int *kk,
                            double *hist) {
  if (!h) return fail(MDX_E_ARG, "mdx_get_nose: NULL handle");
  Ctx &c = h->c;
  if (!c.mass.p) return fail(MDX_E_ARG, "mdx_get_nose: system has no masses");
  MDX_HIP(hipSetDevice(c.device));
  double v[4];
  MDX_HIP(hipMemcpyAsync(v, c.nose.p, sizeof(v), hipMemcpyDeviceToHost, c.stream));
  if (hist && c.trvp_k > 0)
    MDX_HIP(hipMemcpyAsync(hist, c.trvp_hist.p, sizeof(double) * c.trvp_k * (1 + 3 * (size_t)c.ns),
                           hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipStreamSynchronize(c.stream));
  if (logs) *logs = v[0];
  if (logs1) *logs1 = v[1];
  if (k) *k = std::max(c.trvp_k, 0);
  if (kk) *kk = c.trvp_kk;
  return 0;
}

extern "C" int mdx_set_npt(mdx_ctx_t *h, double logs, double logs1, double lambda1, int k, int kk,
                           const double *hist) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_npt: NULL handle");
  Ctx &c = h->c;
  if (!c.mass.p) return fail(MDX_E_ARG, "mdx_set_npt: system has no masses");
  if (k < 0 || k > 9 || kk < 0 || kk > k) return fail(MDX_E_ARG, "NPT: need 0 <= kk <= k <= 9");
  MDX_HIP(hipSetDevice(c.device));
  const size_t row = 1 + 3 * (size_t)c.ns;
  std::vector<double> dev;
  if (hist && k > 0) {
    // the device rows: a logs slot (unused by NPT) and the sites
    dev.resize(row * k);
    for (int j = 0; j < k; j++) {
      dev[j * row] = hist[j * (row + 1)];
      std::memcpy(&dev[j * row + 1], hist + j * (row + 1) + 2, sizeof(double) * 3 * c.ns);
    }
  }
  int rc;
  if ((rc = set_trvp(c, k, kk, dev.empty() ? nullptr : dev.data()))) return rc;
  c.npt_logs = logs;
  c.npt_logs1 = logs1;
  c.npt_lam1 = lambda1;
  for (int j = 0; j < 9; j++) {
    c.npt_hlogs[j] = hist && j < k ? hist[j * (row + 1)] : 0.0;
    c.npt_hlam[j] = hist && j < k ? hist[j * (row + 1) + 1] : 0.0;
  }
  return 0;
}

extern "C" int mdx_get_npt(mdx_ctx_t *h, double *logs, double *logs1, double *lambda1, int *k,
                           int *kk, double *hist) {
  if (!h) return fail(MDX_E_ARG, "mdx_get_npt: NULL handle");
  Ctx &c = h->c;
  if (!c.mass.p) return fail(MDX_E_ARG, "mdx_get_npt: system has no masses");
  const int kn = std::max(c.trvp_k, 0);
  if (hist && kn > 0) {
    MDX_HIP(hipSetDevice(c.device));
    const size_t row = 1 + 3 * (size_t)c.ns;
    std::vector<double> dev(row * kn);
    MDX_HIP(hipMemcpyAsync(dev.data(), c.trvp_hist.p, sizeof(double) * dev.size(),
                           hipMemcpyDeviceToHost, c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
    for (int j = 0; j < kn; j++) {
      hist[j * (row + 1)] = c.npt_hlogs[j];
      hist[j * (row + 1) + 1] = c.npt_hlam[j];
      std::memcpy(hist + j * (row + 1) + 2, &dev[j * row + 1], sizeof(double) * 3 * c.ns);
    }
  }
  if (logs) *logs = c.npt_logs;
  if (logs1) *logs1 = c.npt_logs1;
  if (lambda1) *lambda1 = c.npt_lam1;
  if (k) *k = kn;
  if (kk) *kk = c.trvp_kk;
  return 0;
}

namespace {

// isotropic NPT, before the forces (sim/trvpscale.c:43-70 with rescale =
// "xyz"): the predicted h dlogs/dt (Vp0) and h dlambda/dt (Vp1) from the
// current logs1, lambda1 and their history, Vfactor = Vp0 + No.NPT Vp1
// (No.NPT = 1 + 3/No.f for MTK, 1 for Hoover, sim/siminit.c:607), the
// constraint scaling exp(2 lambda1 - lambda1(t-h)) (exp(lambda1) while the
// predictor is shorter than 2), then the history shift of both scalars
struct NptPred {
  double Vp0, Vp1, vfac, bscale;
};

NptPred npt_predict(Ctx &c, const mdx_step_t *p, int kk) {
  const Trvp tv = trvp_coef(kk, c.trvp_k);
  NptPred q{};
  for (int j = 0; j <= kk; j++) {
    q.Vp0 += tv.b[j] * (j ? c.npt_hlogs[j - 1] : c.npt_logs1);
    q.Vp1 += tv.b[j] * (j ? c.npt_hlam[j - 1] : c.npt_lam1);
  }
  const double NPT = 1. + (p->thermostat == 22) * 3. / p->nf;
  q.vfac = q.Vp0 + NPT * q.Vp1;
  q.bscale = kk > 1 ? std::exp(2 * c.npt_lam1 - c.npt_hlam[0]) : std::exp(c.npt_lam1);
  for (int j = tv.knew; j >= 2; j--) {
    c.npt_hlogs[j - 1] = c.npt_hlogs[j - 2];
    c.npt_hlam[j - 1] = c.npt_hlam[j - 2];
  }
  if (tv.knew >= 1) {
    c.npt_hlogs[0] = c.npt_logs1;
    c.npt_hlam[0] = c.npt_lam1;
  }
  return q;
}

// isotropic NPT after SHAKE (sim/thermo.c:58-99, masses No.M_T, No.M_P and
// No.fx of sim/siminit.c:604-626 with No.ncoord = 3): En.Pref at the
// current volume, the leap-frog of logs and of lambda = log L, then the
// sites r(t+h) and the box scaled by exp(lambda(t+h) - lambda(t)).  kin
// is En.kin (doubled); out has this step's measured outputs.
int npt_update(mdx_ctx_t *h, const mdx_step_t *p, const NptPred &q, double kin,
               mdx_step_out_t *out) {
  Ctx &c = h->c;
  const double hh = p->h * p->h, T = p->T, V = c.L[0] * c.L[1] * c.L[2];
  const double invf = (double)(p->thermostat == 22) / p->nf;
  const double fx = p->nf + (p->tau_P < .999e6 ? 1 : 0);
  const double M_T = p->tau * p->tau * T * p->nf, M_Th = M_T / 2;
  const double M_P = p->tau_P * p->tau_P * T * (p->nf + 3) * 3, M_Ph = M_P / 2;
  mdx_pressure_t pr;
  int rc = mdx_pressure(out, c.L, 1.0, p->Ecorr, 0, &pr);
  if (rc) return rc;
  const double Pref = p->virial == 1 ? pr.Pevir_c : pr.Ptr_c;
  const double ldlogs = c.npt_logs1, lam0 = std::log(c.L[0]);
  double ext = fx * T * c.npt_logs;
  const double kinT = M_Th / hh * (q.Vp0 * q.Vp0), kinP = M_Ph / hh * (q.Vp1 * q.Vp1);
  c.npt_logs1 += hh * (kin + 6 * kinP - fx * T) / M_T;
  c.npt_logs += c.npt_logs1;
  ext += kinT + kinP * 3 + V * p->P;
  c.npt_lam1 += hh * (V * (Pref - p->P) + kin * invf) / M_P - q.Vp0 * q.Vp1;
  const double scale = std::exp((lam0 + c.npt_lam1) - lam0);
  out->En_ext = ext;
  out->logs = c.npt_logs;
  out->dlogs = (ldlogs + c.npt_logs1) / (2 * p->h);
  out->Pref = Pref;
  out->V = V;
  out->npt_lambda = lam0;
  out->npt_dlambda = q.Vp1 / p->h;
  k_rescale<<<nblk(3L * c.ns), kBlock, 0, c.stream>>>(3 * c.ns, c.r.p, scale);
  if (hipGetLastError() != hipSuccess) return fail(MDX_E_GPU, "mdx_md_steps: launch");
  const double L[3] = {c.L[0] * scale, c.L[1] * scale, c.L[2] * scale};
  if ((rc = mdx_set_box(h, L))) return rc;
  for (int d = 0; d < 3; d++) out->L[d] = c.L[d];
  return 0;
}

inline bool is_npt(int thermostat) { return thermostat == 22 || thermostat == 23; }

// the launches of one MD step (no host synchronization): for Drudes the
// ASPC prediction with predictor length kk and scf_iter forces()+selffield()
// with mixing omega, then forces (others), SHAKE + leap-frog, the reduction
// and the Berendsen scaling.  NPT: vfac and bscale of this step from the
// host (its thermostat update and rescaling follow on the host)
int launch_step(Ctx &c, const mdx_step_t *p, int measure, int kk, double omega,
                double vfac = 0, double bscale = 1) {
  const int nb = nblk(c.nmol);
  int rc;
  if (c.polar) {
    k_aspc<<<nblk(3L * c.ns), kBlock, 0, c.stream>>>(3 * c.ns, c.r.p + 3 * c.ns, c.r1.p + 3 * c.ns,
                                                     c.aspc_hist.p, c.aspc_k, kk, aspc_coef(kk));
    const int niter = p->scf_iter > 0 ? p->scf_iter : 1;
    for (int it = 0; it < niter; it++) {
      if ((rc = forces_device(c, measure && it == niter - 1))) return rc;
      launch_selffield(c, omega);
    }
  } else if ((rc = forces_device(c, measure))) {
    return rc;
  }
  if (c.timing) c.tm.start(T_SHAKE, c.stream);
  const bool nose = p->thermostat == 2, npt = is_npt(p->thermostat);
  const long stride = 1 + 3L * c.ns;
  const Trvp tv = trvp_coef(nose || npt ? kk : 0, nose || npt ? c.trvp_k : 0);
  k_shake<<<nb, kBlock, 0, c.stream>>>(c.nmol, c.molfirst.p, c.cons_first.p, c.cons_ij.p,
                                       c.cons_bondq.p, c.mol_omega.p, c.mass.p, p->h, p->eps,
                                       p->maxit, measure, c.r.p, c.r1.p, c.f.p, c.pwork.p,
                                       c.mdpart.p, nose ? PRED_NOSE : npt ? PRED_HOST : PRED_NONE,
                                       c.nose.p, vfac, bscale, c.trvp_hist.p, stride, tv,
                                       c.shk_nsp > 1 ? c.shk_sp.p : nullptr, c.shk_nit.p);
  if (c.timing) c.tm.stop(T_SHAKE, c.stream);
  k_md_reduce<<<1, kBlock, 0, c.stream>>>(c.mdpart.p, nb, c.mdacc.p, c.mdflag.p);
  if (p->thermostat == 1)
    k_berendsen<<<nblk(3L * c.ns), kBlock, 0, c.stream>>>(3 * c.ns, c.r1.p, c.mdacc.p, p->h,
                                                          p->T, p->tau, p->nf, c.mdflag.p);
  if (nose)
    k_nose<<<1, 64, 0, c.stream>>>(c.mdacc.p, c.nose.p, c.trvp_hist.p, stride, tv.knew, p->h,
                                   p->T, p->tau, p->nf);
  if (p->thermostat >= 11 && p->thermostat <= 13)
    k_friction<<<nb, kBlock, 0, c.stream>>>(c.nmol, c.molfirst.p, c.mass.p, c.r1.p, c.mdacc.p,
                                            p->h, p->T, p->tau, p->nf, p->nf_tr, p->T_tr_in,
                                            p->thermostat);
  if (hipGetLastError() != hipSuccess) return fail(MDX_E_GPU, "mdx_md_steps: launch");
  return 0;
}

// the step parameters a captured graph depends on
bool same_step(const mdx_step_t &a, const mdx_step_t &b) {
  return a.h == b.h && a.eps == b.eps && a.maxit == b.maxit && a.thermostat == b.thermostat &&
         a.T == b.T && a.tau == b.tau && a.nf == b.nf && a.aspc_k == b.aspc_k &&
         a.scf_iter == b.scf_iter && a.scf_omega == b.scf_omega && a.trvp_k == b.trvp_k &&
         a.nf_tr == b.nf_tr && a.T_tr_in == b.T_tr_in && a.P == b.P && a.tau_P == b.tau_P &&
         a.Ecorr == b.Ecorr && a.virial == b.virial;
}

// an unmeasured step as a hipGraph, captured once per step parameters and
// engine configuration (buffers, grid, k-space split: c.gen) and replayed
int graph_step(Ctx &c, const mdx_step_t *p, int kk, double omega) {
  if (!c.graph_ok || c.graph_gen != c.gen || !same_step(c.graph_p, *p) || c.graph_kk != kk) {
    if (c.graph_exec) { (void)hipGraphExecDestroy(c.graph_exec); c.graph_exec = nullptr; }
    if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
    c.graph_ok = false;
    MDX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(c, p, 0, kk, omega);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess || !g) return fail(MDX_E_GPU, "mdx_md_steps: graph capture failed");
    c.graph = g;
    if (hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0) != hipSuccess)
      return fail(MDX_E_GPU, "mdx_md_steps: graph instantiation failed");
    c.graph_ok = true;
    c.graph_gen = c.gen;
    c.graph_p = *p;
    c.graph_kk = kk;
  }
  MDX_HIP(hipGraphLaunch(c.graph_exec, c.stream));
  return 0;
}

}  // namespace

extern "C" int mdx_set_graph(mdx_ctx_t *h, int on) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_graph: NULL");
  h->c.use_graph = on != 0;
  return 0;
}

extern "C" int mdx_set_overlap(mdx_ctx_t *h, int on) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_overlap: NULL");
  MDX_HIP(hipSetDevice(h->c.device));
  MDX_HIP(hipStreamSynchronize(h->c.stream));
  h->c.overlap = on != 0;
  h->c.gen++;                          // a captured step holds the old shape
  return 0;
}

extern "C" int mdx_md_steps(mdx_ctx_t *h, const mdx_step_t *p, int nsteps, mdx_step_out_t *out) {
  if (!h || !p || !out || nsteps < 0) return fail(MDX_E_ARG, "mdx_md_steps: bad argument");
  Ctx &c = h->c;
  if (!c.state_on_device) return fail(MDX_E_ARG, "mdx_md_steps: upload the state first");
  if (!(p->h > 0) || !(p->eps > 0)) return fail(MDX_E_ARG, "mdx_md_steps: h and eps must be > 0");
  const int th = p->thermostat;
  const bool npt = is_npt(th);
  if (!(th >= 0 && th <= 2) && !(th >= 11 && th <= 13) && !npt)
    return fail(MDX_E_ARG, "mdx_md_steps: thermostat 0 (none), 1 (Berendsen), 2 (Nose-Hoover), "
                           "11-13 (friction on translations, intramolecular, both) or 22, 23 "
                           "(MTK, Hoover NPT)");
  if (npt && (c.polar || !(p->tau_P > 0) || !std::isfinite(p->P) || !std::isfinite(p->Ecorr) ||
              !(p->virial == 0 || p->virial == 1 || p->virial == 3)))
    return fail(MDX_E_ARG, "mdx_md_steps: NPT needs tau_P > 0, a finite P and Ecorr, virial 0, "
                           "1 or 3, and no Drude charges");
  if (th >= 11 && th <= 13 && (c.polar || !(p->nf_tr > 0 && p->nf - p->nf_tr > 0) ||
                   (th == 13 && !(p->T_tr_in > 0))))
    return fail(MDX_E_ARG, "mdx_md_steps: the friction thermostats need 0 < nf_tr < nf (and "
                           "T_tr_in > 0 for 13), not for Drude systems");
  if (p->thermostat && !(p->tau > 0 && p->nf > 0 && p->T > 0))
    return fail(MDX_E_ARG, "mdx_md_steps: the thermostat needs T, tau, nf > 0");
  if (p->thermostat == 2 && c.polar)
    return fail(MDX_E_ARG, "mdx_md_steps: Nose-Hoover is not carried for Drude systems");
  if ((p->thermostat == 2 || npt) && (p->trvp_k < 0 || p->trvp_k > 9))
    return fail(MDX_E_ARG, "mdx_md_steps: trvp_k in 0..9");
  if (c.polar && !(p->scf_omega <= 1))
    return fail(MDX_E_ARG, "mdx_md_steps: scf_omega must be <= 1 (<= 0: stability value)");
  MDX_HIP(hipSetDevice(c.device));
  int rc;
  if (c.polar && c.aspc_k != p->aspc_k && (rc = set_aspc(c, p->aspc_k, 0, nullptr))) return rc;
  if ((p->thermostat == 2 || npt) && c.trvp_k != p->trvp_k) {
    if ((rc = set_trvp(c, p->trvp_k, 0, nullptr))) return rc;
    std::fill(c.npt_hlogs, c.npt_hlogs + 9, 0.0);
    std::fill(c.npt_hlam, c.npt_hlam + 9, 0.0);
  }
  std::memset(out, 0, sizeof(*out));
  out->lambda = 1;
  const int every = p->measure_every > 0 ? p->measure_every : nsteps;
  double acc[ACC_N], md[MD_N], sf[SF_N], flag[4], nv[4];
  reset_flags(c);
  rc = 0;
  for (int s = 1; s <= nsteps && !rc; s++) {
    const int measure = npt || (s % every == 0) || s == nsteps;
    // predictor length of this step (ASPC: grows to aspc_k, sim/pred2.c:
    // 67-77; Nose-Hoover and NPT: grows to trvp_k, sim/trvpscale.c:95)
    const bool nose = p->thermostat == 2 || npt;
    const int kk = nose ? c.trvp_kk : c.aspc_kk;
    if (c.polar && c.aspc_kk < c.aspc_k) c.aspc_kk++;
    if (nose && c.trvp_kk < c.trvp_k) c.trvp_kk++;
    const NptPred np = npt ? npt_predict(c, p, kk) : NptPred{0, 0, 0, 1};
    const double omega = !c.polar ? 0.0
                         : p->scf_omega > 0 ? p->scf_omega
                                            : (c.aspc_kk + 2) / (double)(2 * c.aspc_kk + 3);
    // unmeasured steps of the steady state replay a captured graph
    const bool steady = nose ? kk == c.trvp_k : (!c.polar || kk == c.aspc_k);
    if (!measure && steady && c.use_graph && !c.timing) rc = graph_step(c, p, kk, omega);
    else rc = launch_step(c, p, measure, kk, omega, np.vfac, np.bscale);
    if (rc) break;
    if (measure) {
      if (hipMemcpyAsync(acc, c.acc.p, sizeof(acc), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
          hipMemcpyAsync(md, c.mdacc.p, sizeof(md), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
          hipMemcpyAsync(flag, c.mdflag.p, sizeof(flag), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
          (th == 2 && hipMemcpyAsync(nv, c.nose.p, sizeof(nv), hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
          (c.polar && hipMemcpyAsync(sf, c.sfacc.p, sizeof(sf), hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
          hipStreamSynchronize(c.stream) != hipSuccess) {
        rc = fail(MDX_E_GPU, "mdx_md_steps: download");
        break;
      }
      if (flag[1] > 0) { rc = fail(MDX_E_NOCONV, "SHAKE: too many iterations"); break; }
      if (flag[3] > 0) {
        rc = fail(MDX_E_RANGE, "selffield: Drude farther than el.minqq=1 A from its centre");
        break;
      }
      const double hh = p->h * p->h, hscale = 0.5 / hh;
      fill_out(c, acc, 1, &out->frc);
      out->En_kin = md[MD_EKIN] * hscale / 2;
      out->En_kin_tr = md[MD_KTR] * hscale / 2;
      for (int k = 0; k < 6; k++) {
        out->Pkin[k] = md[MD_PKIN + k] * hscale;
        out->PTvirc[k] = md[MD_PTV + k] / hh;
      }
      out->virc = md[MD_VIRC] / hh;
      out->shake_iter_max = (int)flag[2];
      out->lambda = flag[0];
      if (th == 2) {
        out->logs = nv[0];
        out->En_ext = nv[2];
        out->dlogs = nv[3];
      }
      if (c.polar) {
        out->frc.En_el += sf[SF_EEL];
        for (int k = 0; k < 6; k++) out->frc.Pvir[k] += sf[SF_PV + k];
        out->En_self = sf[SF_SELF] / 2;
        out->scf_maxerr = sqrt(sf[SF_ERR]);
      }
      if (npt && (rc = npt_update(h, p, np, md[MD_EKIN] * hscale, out))) break;
    }
  }
  if (!rc) {
    unsigned long long nit[64];
    if (hipMemcpyAsync(nit, c.shk_nit.p, sizeof(unsigned long long) * c.shk_nsp,
                       hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
      return fail(MDX_E_GPU, "mdx_md_steps: download");
    for (int k = 0; k < c.shk_nsp; k++) out->shake_nit += (long long)nit[k];
  }
  (void)hipStreamSynchronize(c.stream);
  return rc;
}

extern "C" int mdx_set_shake_species(mdx_ctx_t *h, const int *msp, int nspec) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_shake_species: NULL");
  Ctx &c = h->c;
  if (!msp) nspec = 1;
  if (nspec < 1 || nspec > 64) return fail(MDX_E_ARG, "mdx_set_shake_species: nspec in 1..64");
  std::vector<int> v(c.nmol, 0);
  if (msp)
    for (int m = 0; m < c.nmol; m++) {
      if (msp[m] < 0 || msp[m] >= nspec)
        return fail(MDX_E_ARG, "mdx_set_shake_species: species index out of [0, nspec)");
      v[m] = msp[m];
    }
  MDX_HIP(hipSetDevice(c.device));
  MDX_HIP(hipStreamSynchronize(c.stream));
  int rc;
  if ((rc = c.shk_sp.alloc(std::max(1, c.nmol)))) return rc;
  if (c.nmol)
    MDX_HIP(hipMemcpy(c.shk_sp.p, v.data(), sizeof(int) * c.nmol, hipMemcpyHostToDevice));
  c.shk_sp_host = msp ? v : std::vector<int>();
  c.shk_nsp = nspec;
  c.gen++;                             // captured steps hold the species pointer
  return 0;
}

extern "C" int mdx_set_shake_omega(mdx_ctx_t *h, const double *omega) {
  if (!h || !omega) return fail(MDX_E_ARG, "mdx_set_shake_omega: NULL");
  Ctx &c = h->c;
  for (int k = 0; k < c.shk_nsp; k++)
    if (!(omega[k] > 0 && omega[k] < 2))
      return fail(MDX_E_ARG, "mdx_set_shake_omega: omega in (0, 2)");
  if (!c.nmol) return 0;
  std::vector<double> om(c.nmol);
  for (int m = 0; m < c.nmol; m++) om[m] = omega[c.shk_sp_host.empty() ? 0 : c.shk_sp_host[m]];
  MDX_HIP(hipSetDevice(c.device));
  MDX_HIP(hipStreamSynchronize(c.stream));
  // the same buffer as before: captured steps read the new values
  MDX_HIP(hipMemcpy(c.mol_omega.p, om.data(), sizeof(double) * c.nmol, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mdx_get_shake_nit(mdx_ctx_t *h, long long *nit, int nspec) {
  if (!h || !nit) return fail(MDX_E_ARG, "mdx_get_shake_nit: NULL");
  Ctx &c = h->c;
  if (nspec != c.shk_nsp) return fail(MDX_E_ARG, "mdx_get_shake_nit: nspec differs from the species set");
  MDX_HIP(hipSetDevice(c.device));
  MDX_HIP(hipStreamSynchronize(c.stream));
  unsigned long long v[64];
  MDX_HIP(hipMemcpy(v, c.shk_nit.p, sizeof(unsigned long long) * nspec, hipMemcpyDeviceToHost));
  for (int k = 0; k < nspec; k++) nit[k] = (long long)v[k];
  return 0;
}
