// Adjoint of the Ewald electrostatics with respect to the site charges, for
// fitting charges to forces or energies (FP64 in either mode of the
// context; non-polarizable systems).
//
// For weights w_i (dL/dF_i of a loss of the forces) and G = sum_i w_i . F_i,
// the engine returns per charged site m
//   gF_m = dG/dq_m  (w held fixed),   gE_m = dE_el/dq_m,
// E_el the electrostatic energy of the force call (r-space, exclusion
// corrections, k-space, self, background, diagonal and cell-dipole terms,
// sim/ewald.c:983-997 with cook/ljqqnp.c:104-206).  Every term of the
// forces is q_a q_b times a function of the positions, so
//   r-space pair (a,b), f_a = q_a q_b erd (r_a - r_b) = -f_b:
//     gF_a += q_b erd (w_a - w_b).(r_a - r_b),  gE_a += q_b u
//   (u, erd from cook's hyperbolic spline of erfc(alpha r)/r, or for EXCL
//   pairs of a molecule the exact -erf(alpha r)/r, sim/erexact.c:25-53);
//   k-space E_k = sum_H c_k |S(k)|^2, S(k) = sum_j q_j e^{i K.r_j}, over
//   the half-space list H of the force call (triples kx, ky, kz >= 0 and
//   their sign variants (kx, +-ky, +-kz), c_k = w e(k) / 4 pi with the
//   reference's ekk e(k) = 4 exp(-pi^2 k^2/alpha^2)/(k^2 V) and w =
//   2^(#nonzero components)/8, K = 2 pi (kx/Lx, ky/Ly, kz/Lz)):
//     gF_m = -2 sum_H c_k Re[e^{-iK.r_m} W(k) + i (K.w_m) e^{iK.r_m} S*(k)],
//     gE_m =  2 sum_H c_k Re[e^{-iK.r_m} S(k)],
//     W(k) = i sum_j (K.w_j) q_j e^{iK.r_j};
//   self -alpha/sqrt(pi) sum q^2, background -pi/2 (Q/alpha)^2 / V, Ediag
//   (proportional to sum q^2) and the cell-dipole energy sum_d ce_d/2 M_d^2
//   with the field -cm M + cb r on each charge (sim/ewald.c:865-931) are
//   differentiated on the host.
// The pair terms are recomputed on the cell grid of the force call (one
// site per lane, full shell, the stencil's x slices in separate
// workgroups); the k-space sums S, U run per triple (4 variants at once)
// over charge splits, with the phases e^{2 pi i k x/L} of a tile of
// charges tabulated in LDS, and the per-charge sums over the triples
// take the kz phases by recursion along each (kx, ky) row.  Partial sums
// are added in a fixed order, so a call is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mdx_devfn.h"
#include "mdx_internal.h"

using namespace mdx;

namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;
constexpr int kRec = 24;   // doubles per triple record

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

// r-space: per site s (cell order) the sums over its partners inside the
// cutoff in the x slice dx = blockIdx.y - w of the stencil;
// rF/rE[blockIdx.y][i] for the original index i
__global__ __launch_bounds__(kBlock) void k_qadj_r(
    int ns, int nx, int ny, int nz, double Lx, double Ly, double Lz, int w, double cq,
    const double4 *__restrict__ spos, const int4 *__restrict__ sinfo,
    const int *__restrict__ scell, const int *__restrict__ cstart, const double *__restrict__ erf,
    double sgrid, double alphar, const double *__restrict__ wgt, double *__restrict__ rF,
    double *__restrict__ rE) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= ns) return;
  const double4 pi = spos[s];
  const int4 ii = sinfo[s];
  double g = 0, p = 0;
  if (pi.w != 0) {
    const double wx = wgt[3 * ii.w], wy = wgt[3 * ii.w + 1], wz = wgt[3 * ii.w + 2];
    const int c = scell[s];
    const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
    const int dx = (int)blockIdx.y - w;
    int ix = cx + dx;
    double sx = 0;
    if (ix < 0) { ix += nx; sx = -Lx; } else if (ix >= nx) { ix -= nx; sx = Lx; }
    for (int dy = -w; dy <= w; dy++) {
      int iy = cy + dy;
      double sy = 0;
      if (iy < 0) { iy += ny; sy = -Ly; } else if (iy >= ny) { iy -= ny; sy = Ly; }
      for (int dz = -w; dz <= w; dz++) {
        int iz = cz + dz;
        double sz = 0;
        if (iz < 0) { iz += nz; sz = -Lz; } else if (iz >= nz) { iz -= nz; sz = Lz; }
        const int nc = (ix * ny + iy) * nz + iz;
        for (int j = cstart[nc]; j < cstart[nc + 1]; j++) {
          if (j == s) continue;
          const double4 pj = spos[j];
          if (pj.w == 0) continue;
          const double ddx = pi.x - (pj.x + sx), ddy = pi.y - (pj.y + sy), ddz = pi.z - (pj.z + sz);
          const double rr = ddx * ddx + ddy * ddy + ddz * ddz;
          if (rr >= cq) continue;
          const int4 jj = sinfo[j];
          double u, erd;
          if (jj.y == ii.y && ((ii.z >> (jj.x >> 16)) & 1)) {
            u = exact_erud(rr, alphar, erd);              // EXCL correction
          } else {
            const double *t = erf + 8 * (int)(rr * sgrid);
            erd = t[4] / (t[5] + rr) + t[3];
            u = t[1] / (t[2] + rr) + t[0];
          }
          const int jo = jj.w;
          const double dw = (wx - wgt[3 * jo]) * ddx + (wy - wgt[3 * jo + 1]) * ddy +
                            (wz - wgt[3 * jo + 2]) * ddz;
          g += pj.w * erd * dw;
          p += pj.w * u;
        }
      }
    }
  }
  rF[(size_t)blockIdx.y * ns + ii.w] = g;
  rE[(size_t)blockIdx.y * ns + ii.w] = p;
}

// the charge list as (x/Lx, y/Ly, z/Lz, q) and w
__global__ void k_qadj_gather(int nq, const int *__restrict__ qsite, const double *__restrict__ r,
                              const double *__restrict__ charge, const double *__restrict__ wgt,
                              double iLx, double iLy, double iLz, double4 *__restrict__ qa,
                              double4 *__restrict__ qb) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= nq) return;
  const int k = qsite[j];
  const double q = charge[k];
  qa[j] = make_double4(r[3 * k] * iLx, r[3 * k + 1] * iLy, r[3 * k + 2] * iLz, q);
  qb[j] = make_double4(wgt[3 * k], wgt[3 * k + 1], wgt[3 * k + 2], 0);
}

// e^{i(a + s b + t c)} for the variants (s,t) = (+,+), (+,-), (-,+), (-,-)
// from e^{i(a+b)} = (pr, pi), e^{i(a-b)} = (mr, mi) and e^{ic} = (cc, sc)
__device__ inline void variants(double pr, double pi, double mr, double mi, double cc, double sc,
                                double (&C)[4], double (&S)[4]) {
  C[0] = pr * cc - pi * sc; S[0] = pi * cc + pr * sc;
  C[1] = pr * cc + pi * sc; S[1] = pi * cc - pr * sc;
  C[2] = mr * cc - mi * sc; S[2] = mi * cc + mr * sc;
  C[3] = mr * cc + mi * sc; S[3] = mi * cc - mr * sc;
}

// k-space sums per triple over the charges [j0, j1) of split blockIdx.y:
// for each variant v, S_v = sum q e^{iK_v.r} and U_v = sum (K_v.q w)
// e^{iK_v.r} (W = i U); part[(split * ntrip + t) * 16 + 4 v + (Sr, Si, Ur, Ui)].
// Per tile of jt charges the phases e^{2 pi i k f} of k = 0..kxmax (f = x/Lx),
// 0..kymax (y/Ly), 0..kzmax (z/Lz) are tabulated in LDS ([charge][ne]
// complex, ne = kxmax + kymax + kzmax + 3), then every triple reads its three.
__global__ __launch_bounds__(kBlock) void k_qadj_sfac(int ntrip, int nq, int nsplit, int nx1,
                                                      int ny1, int ne, int jt,
                                                      const double *__restrict__ tab,
                                                      const double4 *__restrict__ qa,
                                                      const double4 *__restrict__ qb,
                                                      double *__restrict__ part) {
  extern __shared__ double2 ph[];          // [jt][ne], then qb of the tile
  double4 *lb = (double4 *)(ph + (size_t)jt * ne);
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int split = blockIdx.y;
  const int j0 = (int)((long)nq * split / nsplit), j1 = (int)((long)nq * (split + 1) / nsplit);
  int ix = 0, iy = 0, iz = 0;
  double Kx = 0, Ky = 0, Kz = 0;
  if (t < ntrip) {
    const double *rec = tab + (size_t)t * kRec;
    ix = (int)rec[0]; iy = nx1 + (int)rec[1]; iz = nx1 + ny1 + (int)rec[2];
    Kx = rec[4]; Ky = rec[5]; Kz = rec[6];
  }
  double acc[16];
  for (int k = 0; k < 16; k++) acc[k] = 0;
  for (int jb = j0; jb < j1; jb += jt) {
    const int n = min(jt, j1 - jb);
    __syncthreads();
    for (int e = threadIdx.x; e < n * ne; e += kBlock) {
      const int j = e / ne, i = e - j * ne;
      const double4 a = qa[jb + j];
      double f;
      int k;
      if (i < nx1) { k = i; f = a.x; }
      else if (i < nx1 + ny1) { k = i - nx1; f = a.y; }
      else { k = i - nx1 - ny1; f = a.z; }
      double cs, sn;
      cs2pi_d(k * f, cs, sn);
      ph[e] = make_double2(i < nx1 ? a.w * cs : cs, i < nx1 ? a.w * sn : sn);
    }
    if (threadIdx.x < n) lb[threadIdx.x] = qb[jb + threadIdx.x];
    __syncthreads();
    if (t < ntrip) {
      for (int j = 0; j < n; j++) {
        // the x phase carries q, so C, S are q e^{iK.r}
        const double2 X = ph[j * ne + ix], Y = ph[j * ne + iy], Z = ph[j * ne + iz];
        const double4 b = lb[j];
        const double pr = X.x * Y.x - X.y * Y.y, pi = X.y * Y.x + X.x * Y.y;
        const double mr = X.x * Y.x + X.y * Y.y, mi = X.y * Y.x - X.x * Y.y;
        double C[4], S[4];
        variants(pr, pi, mr, mi, Z.x, Z.y, C, S);
        const double kwx = Kx * b.x, kwy = Ky * b.y, kwz = Kz * b.z;
        const double kw[4] = {kwx + kwy + kwz, kwx + kwy - kwz, kwx - kwy + kwz, kwx - kwy - kwz};
#pragma unroll
        for (int v = 0; v < 4; v++) {
          acc[4 * v] += C[v];
          acc[4 * v + 1] += S[v];
          acc[4 * v + 2] += kw[v] * C[v];
          acc[4 * v + 3] += kw[v] * S[v];
        }
      }
    }
  }
  if (t < ntrip) {
    double *o = part + ((size_t)split * ntrip + t) * 16;
    for (int k = 0; k < 16; k++) o[k] = acc[k];
  }
}

// the split sums in a fixed order into the records' S/U slots
__global__ void k_qadj_ksum(int ntrip, int nsplit, const double *__restrict__ part,
                            double *__restrict__ tab) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ntrip * 16) return;
  const int t = i / 16, k = i % 16;
  double s = 0;
  for (int p = 0; p < nsplit; p++) s += part[((size_t)p * ntrip + t) * 16 + k];
  tab[(size_t)t * kRec + 8 + k] = s;
}

// per charge j, over the (kx, ky) rows [r0, r1) of split blockIdx.y: the
// k-space parts of gF and gE (row and triple records read alike by every
// lane); e^{i kz c} by recursion along the row from e^{ic}, e^{i kx a} and
// e^{i ky b} directly per row.  kF/kE[blockIdx.y][j].
__global__ __launch_bounds__(kBlock) void k_qadj_kfield(int nrows, int nq, int nsplit,
                                                        const int4 *__restrict__ rows,
                                                        const double *__restrict__ tab,
                                                        const int *__restrict__ qsite,
                                                        const double4 *__restrict__ qa,
                                                        const double *__restrict__ wgt,
                                                        double *__restrict__ kF,
                                                        double *__restrict__ kE) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= nq) return;
  const int r0 = (int)((long)nrows * blockIdx.y / nsplit);
  const int r1 = (int)((long)nrows * (blockIdx.y + 1) / nsplit);
  const double4 a = qa[j];
  const int site = qsite[j];
  const double wx = wgt[3 * site], wy = wgt[3 * site + 1], wz = wgt[3 * site + 2];
  double c1, s1;
  cs2pi_d(a.z, c1, s1);
  double g = 0, e = 0;
  int kxp = -1;
  double ex = 1, sx = 0;
  for (int r = r0; r < r1; r++) {
    const int4 row = rows[r];
    if (row.x != kxp) {
      cs2pi_d(row.x * a.x, ex, sx);
      kxp = row.x;
    }
    double ey, sy;
    cs2pi_d(row.y * a.y, ey, sy);
    const double pr = ex * ey - sx * sy, pi = sx * ey + ex * sy;
    const double mr = ex * ey + sx * sy, mi = sx * ey - ex * sy;
    double cz = 1, sz = 0;
    for (int kz = 0; kz <= row.z; kz++) {
      if (kz) {
        const double t = cz * c1 - sz * s1;
        sz = sz * c1 + cz * s1;
        cz = t;
      }
      const int t = row.w + kz;
      if (t < 0) continue;                     // (0,0,0)
      const double *rec = tab + (size_t)t * kRec;
      double C[4], S[4];
      variants(pr, pi, mr, mi, cz, sz, C, S);
      const double c2 = 2 * rec[3];
      const double kwx = rec[4] * wx, kwy = rec[5] * wy, kwz = rec[6] * wz;
      const double kw[4] = {kwx + kwy + kwz, kwx + kwy - kwz, kwx - kwy + kwz, kwx - kwy - kwz};
      double gt = 0, et = 0;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const double Sr = rec[8 + 4 * v], Si = rec[9 + 4 * v];
        const double Ur = rec[10 + 4 * v], Ui = rec[11 + 4 * v];
        // Re[e^{-iKr} W] with W = i U, and Re[i (K.w) e^{iKr} S*]
        gt += (S[v] * Ur - C[v] * Ui) + kw[v] * (Si * C[v] - Sr * S[v]);
        et += C[v] * Sr + S[v] * Si;
      }
      g -= c2 * gt;
      e += c2 * et;
    }
  }
  kF[(size_t)blockIdx.y * nq + j] = g;
  kE[(size_t)blockIdx.y * nq + j] = e;
}

// per charge j: the r-space slices and k-space splits summed in a fixed
// order into gF/gE of its site
__global__ void k_qadj_final(int nq, int ns, int nr, int nk, const int *__restrict__ qsite,
                             const double *__restrict__ rF, const double *__restrict__ rE,
                             const double *__restrict__ kF, const double *__restrict__ kE,
                             double *__restrict__ gF, double *__restrict__ gE) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= nq) return;
  const int site = qsite[j];
  double f = 0, e = 0;
  for (int y = 0; y < nr; y++) {
    f += rF[(size_t)y * ns + site];
    e += rE[(size_t)y * ns + site];
  }
  for (int y = 0; y < nk; y++) {
    f += kF[(size_t)y * nq + j];
    e += kE[(size_t)y * nq + j];
  }
  gF[site] = f;
  gE[site] = e;
}

// the triple records of the context's k-set (rows of kx, ky with kz =
// 0..kzmax, (0,0,0) left out): kx, ky, kz, c = w e / 4 pi, K = 2 pi (kx/Lx,
// ky/Ly, kz/Lz), pad, then 16 slots for S/U; and per row (kx, ky, kzmax,
// index of its kz = 0 triple)
void triple_records(const Ctx &c, std::vector<double> &tab, std::vector<int> &rows) {
  const KLayout &K = c.K;
  const double V = c.L[0] * c.L[1] * c.L[2];
  const double PIaq = -(kPi / c.alpha) * (kPi / c.alpha);
  tab.clear();
  rows.clear();
  int nt = 0;
  for (int row = 0; row < (int)K.row_kx.size(); row++) {
    const int kx = K.row_kx[row], ky = K.row_ky[row];
    const int first = (kx == 0 && ky == 0) ? 1 : 0;
    const int rr[4] = {kx, ky, K.row_kzmax[row], nt - first};
    rows.insert(rows.end(), rr, rr + 4);
    for (int kz = first; kz <= K.row_kzmax[row]; kz++, nt++) {
      const int nz = (kx > 0) + (ky > 0) + (kz > 0);
      const double w = (double)(1 << (nz - 1)) * 0.25;
      const double kxl = kx / c.L[0], kyl = ky / c.L[1], kzl = kz / c.L[2];
      const double kk = kxl * kxl + kyl * kyl + kzl * kzl;
      const double e = 4 / (kk * V) * std::exp(PIaq * kk);
      const double rec[8] = {(double)kx, (double)ky, (double)kz, w * e / (4 * kPi),
                             2 * kPi * kxl, 2 * kPi * kyl, 2 * kPi * kzl, 0};
      tab.insert(tab.end(), rec, rec + 8);
      tab.insert(tab.end(), 16, 0.0);
    }
  }
}

}  // namespace

extern "C" int mdx_adjoint_charges(mdx_ctx_t *h, const double *r, const double *w, double *gF,
                                   double *gE) {
  if (!h || !r || !w || !gF || !gE) return fail(MDX_E_ARG, "mdx_adjoint_charges: NULL argument");
  Ctx &c = h->c;
  if (c.polar)
    return fail(MDX_E_ARG, "mdx_adjoint_charges: systems with Drude charges are not supported");
  MDX_HIP(hipSetDevice(c.device));
  MDX_HIP(hipStreamSynchronize(c.stream));
  const int ns = c.ns, nq = c.nq;
  const int nr = 2 * c.wst + 1;        // r-space x slices
  const KLayout &K = c.K;
  int rc;
  // triple and row records (per context) and the split partials
  if (c.qadj_ntrip == 0 && nq) {
    std::vector<double> tab;
    std::vector<int> rows;
    triple_records(c, tab, rows);
    c.qadj_ntrip = (int)(tab.size() / kRec);
    c.qadj_nrows = (int)(rows.size() / 4);
    if ((rc = c.qadj_tab.alloc(tab.size())) || (rc = c.qadj_rows.alloc(rows.size()))) return rc;
    MDX_HIP(hipMemcpy(c.qadj_tab.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.qadj_rows.p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice));
    // pass 1: about 2^19 threads, at least 64 charges per split; pass 2:
    // about 16 waves per CU
    c.qadj_nsplit = std::max(1, std::min({(1 << 19) / std::max(c.qadj_ntrip, 1), 256, nq / 64}));
    c.qadj_ksplit = std::max(1, std::min(c.qadj_nrows, 4096 / nblk(nq)));
    if ((rc = c.qadj_part.alloc((size_t)c.qadj_nsplit * c.qadj_ntrip * 16))) return rc;
  }
  const int ksplit = nq ? c.qadj_ksplit : 0;
  // pass-1 LDS: phase tables of jt charges (<= 48 KB) and their weights
  const int nx1 = K.kxmax + 1;
  int ny1 = 1, nz1 = 1;
  for (size_t i = 0; i < K.row_ky.size(); i++) {
    ny1 = std::max(ny1, K.row_ky[i] + 1);
    nz1 = std::max(nz1, K.row_kzmax[i] + 1);
  }
  const int ne = nx1 + ny1 + nz1;
  int jt = 64;
  while (jt > 1 && (size_t)jt * (ne * 16 + 32) > 48 * 1024) jt /= 2;
  // per call: weights [3 ns], r-space slices [2 nr ns], charge list [8 nq],
  // k-space splits [2 ksplit nq], results [2 ns]
  // (the charge list starts on a 32-byte boundary)
  const size_t n0 = (3 * (size_t)ns + 2 * (size_t)nr * ns + 3) / 4 * 4;
  const size_t nbuf = n0 + 8 * (size_t)nq + 2 * (size_t)ksplit * nq + 2 * (size_t)ns;
  if ((rc = c.qadj_buf.alloc(nbuf))) return rc;
  double *wd = c.qadj_buf.p, *rF = wd + 3 * (size_t)ns, *rE = rF + (size_t)nr * ns;
  double4 *qa = (double4 *)(wd + n0), *qb = qa + nq;
  double *kF = (double *)(qb + nq), *kE = kF + (size_t)ksplit * nq;
  double *oF = kE + (size_t)ksplit * nq, *oE = oF + ns;
  c.state_on_device = false;           // the device positions are now r
  MDX_HIP(hipMemcpyAsync(c.r.p, r, sizeof(double) * 3 * ns, hipMemcpyHostToDevice, c.stream));
  MDX_HIP(hipMemcpyAsync(wd, w, sizeof(double) * 3 * ns, hipMemcpyHostToDevice, c.stream));
  MDX_HIP(hipMemsetAsync(oF, 0, sizeof(double) * 2 * ns, c.stream));
  if ((rc = bin_device(c))) return rc;
  k_qadj_r<<<dim3(nblk(ns), nr), kBlock, 0, c.stream>>>(
      ns, c.nc[0], c.nc[1], c.nc[2], c.L[0], c.L[1], c.L[2], c.wst, c.cutoff * c.cutoff,
      (const double4 *)c.spos.p, (const int4 *)c.sinfo.p, c.scan_tmp.p, c.cstart.p, c.erf64.p,
      c.erf_sgrid, c.alphar, wd, rF, rE);
  MDX_HIP(hipGetLastError());
  if (nq) {
    k_qadj_gather<<<nblk(nq), kBlock, 0, c.stream>>>(nq, c.q_site.p, c.r.p, c.charge.p, wd,
                                                     1 / c.L[0], 1 / c.L[1], 1 / c.L[2], qa, qb);
    const size_t lds = (size_t)jt * ne * sizeof(double2) + (size_t)jt * sizeof(double4);
    k_qadj_sfac<<<dim3(nblk(c.qadj_ntrip), c.qadj_nsplit), kBlock, lds, c.stream>>>(
        c.qadj_ntrip, nq, c.qadj_nsplit, nx1, ny1, ne, jt, c.qadj_tab.p, qa, qb, c.qadj_part.p);
    k_qadj_ksum<<<nblk((long)c.qadj_ntrip * 16), kBlock, 0, c.stream>>>(
        c.qadj_ntrip, c.qadj_nsplit, c.qadj_part.p, c.qadj_tab.p);
    k_qadj_kfield<<<dim3(nblk(nq), ksplit), kBlock, 0, c.stream>>>(
        c.qadj_nrows, nq, ksplit, (const int4 *)c.qadj_rows.p, c.qadj_tab.p, c.q_site.p, qa, wd,
        kF, kE);
    k_qadj_final<<<nblk(nq), kBlock, 0, c.stream>>>(nq, ns, nr, ksplit, c.q_site.p, rF, rE, kF,
                                                     kE, oF, oE);
    MDX_HIP(hipGetLastError());
  }
  std::vector<double> hF(ns), hE(ns), q(ns);
  MDX_HIP(hipMemcpyAsync(hF.data(), oF, sizeof(double) * ns, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipMemcpyAsync(hE.data(), oE, sizeof(double) * ns, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipMemcpyAsync(q.data(), c.charge.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c.stream));
  MDX_HIP(hipStreamSynchronize(c.stream));
  // host terms: self, background, Ediag, cell dipole with the
  // charged-background field (positions as given, as the force call's M)
  const double V = c.L[0] * c.L[1] * c.L[2];
  double cm[3], ce[3], cb[3];
  ewald_dipole_coefs(c.el_corr, c.el_bg, c.epsinf, c.qnet, V, cm, ce, cb);
  double M[3] = {0, 0, 0}, P[3] = {0, 0, 0}, Qr[3] = {0, 0, 0}, Q = 0;
  for (int i = 0; i < ns; i++) {
    Q += q[i];
    for (int d = 0; d < 3; d++) {
      M[d] += q[i] * r[3 * i + d];
      P[d] += q[i] * w[3 * i + d];
      Qr[d] += q[i] * w[3 * i + d] * r[3 * i + d];
    }
  }
  // dcb/dQ: the background field 4 pi Q r_d / V of el.corr with el.bg
  double dcb[3];
  for (int d = 0; d < 3; d++) dcb[d] = (c.el_corr && (c.el_corr & 1 << d) && c.el_bg) ? 4 * kPi / V : 0;
  const double ediag1 = c.el_diag == 1 ? ewald_ediag(c.L, c.kappa, c.alpha, 1.0) : 0.0;
  for (int m = 0; m < ns; m++) {
    if (q[m] == 0) {
      gF[m] = gE[m] = 0;
      continue;
    }
    double fd = 0, ed = 0;
    for (int d = 0; d < 3; d++) {
      const double rd = r[3 * m + d], wdm = w[3 * m + d];
      fd += wdm * (-cm[d] * M[d] + cb[d] * rd) - cm[d] * rd * P[d] + dcb[d] * Qr[d];
      ed += ce[d] * M[d] * rd;
    }
    gF[m] = hF[m] + fd;
    gE[m] = hE[m] + ed - 2 * c.alphar / std::sqrt(kPi) * q[m] -
            kPi * Q / (c.alpha * c.alpha) / V + 2 * ediag1 * q[m];
  }
  return 0;
}
