// Batched replicas: many independent atomic systems (one site per
// molecule: Lennard-Jones or exp-6 with cook's smooth cutoff, one form for
// the batch, point charges, Ewald,
// leap-frog with Berendsen) of uneven size, run by one launch per kernel
// for all systems, each with its own box, cell grid, spline table, LJ table,
// k-vector set and thermostat; the unmeasured step is captured as one
// hipGraph for the whole batch.
//
// Layout.  Sites of system s occupy [pad_off_s, pad_off_s + ns_s) of the
// padded arrays, pad_off a multiple of 64, so that a wave of the per-site
// MD kernels never holds two systems (padding slots have type -1 and
// charge 0).  The cells of all systems form one counting sort; the sorted
// sites of a system are contiguous.
//
// The physics is the single-system engine's: r-space cook/ljqqnp.c with
// the spline of sim/erfc.c, k-space Ewald(2) of sim/ewald.c in the T
// formulation (mdx_bkspace.hip), leap-frog of sim/shakev3.c and the
// Berendsen scaling of sim/thermo.c.  Sums per system are segmented
// reductions in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mdx_batch.h"

using namespace mdx;
using namespace mdx::batch;

namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;
constexpr int kNE = 11;   // per-site pair sums

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void k_bfill(long n, double *__restrict__ x, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_bzero_int(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__device__ inline double wrapL(double x, double L, double iL) {
  double y = x - L * floor(x * iL);
  return y >= L ? y - L : y;
}

// ---------------------------------------------------------------- binning
__global__ void k_bcell(int npad, const double *__restrict__ r, const int *__restrict__ type,
                        const int *__restrict__ sysid, const DSys *__restrict__ sy,
                        int *__restrict__ cellid, int *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad || type[i] < 0) return;
  const DSys &s = sy[sysid[i]];
  int ic[3];
  for (int d = 0; d < 3; d++) {
    const double x = wrapL(r[3 * i + d], s.L, s.iL);
    const int k = (int)(x * s.nc * s.iL);
    ic[d] = k < 0 ? 0 : (k >= s.nc ? s.nc - 1 : k);
  }
  const int c = s.cell_off + (ic[0] * s.nc + ic[1]) * s.nc + ic[2];
  cellid[i] = c;
  atomicAdd(&cnt[c], 1);
}

// per system (one workgroup each): exclusive scan of its cells' counts,
// starting at the first sorted row of the system (all sites of a system
// fall into its own cells); cursors to zero
__global__ void k_bscan_sys(const DSys *__restrict__ sy, int nsys, int nreal,
                            const int *__restrict__ cnt, int *__restrict__ start,
                            int *__restrict__ cursor) {
  __shared__ int part[kBlock];
  const DSys S = sy[blockIdx.x];
  int base = S.srt_off;
  for (int c0 = 0; c0 < S.ncell; c0 += kBlock) {
    const int c = c0 + threadIdx.x;
    const int v = c < S.ncell ? cnt[S.cell_off + c] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
      const int t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (c < S.ncell) {
      start[S.cell_off + c] = base + part[threadIdx.x] - v;
      cursor[S.cell_off + c] = 0;
    }
    base += part[kBlock - 1];
    __syncthreads();
  }
  if (blockIdx.x == nsys - 1 && threadIdx.x == 0) start[S.cell_off + S.ncell] = nreal;
}

__global__ void k_bscatter(int npad, const int *__restrict__ type, const int *__restrict__ cellid,
                           const int *__restrict__ start, int *__restrict__ cursor,
                           int *__restrict__ sidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad || type[i] < 0) return;
  const int c = cellid[i];
  sidx[start[c] + atomicAdd(&cursor[c], 1)] = i;
}

// sites of each cell in increasing index (deterministic results): a wave
// per cell, bitonic across the lanes up to 64 sites, by rank through tmp
// above
__global__ void k_bsort(int ncell, const int *__restrict__ start, int *__restrict__ sidx,
                        int *__restrict__ tmp) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int b = 0, n = 0;
  if (c < ncell) { b = start[c]; n = start[c + 1] - b; }
  if (n > 1 && n <= 64) {
    int v = lane < n ? sidx[b + lane] : 0x7fffffff;
    for (int k = 2; k <= 64; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int o = __shfl_xor(v, j, 64);
        const bool up = (lane & k) == 0, lower = (lane & j) == 0;
        v = lower == up ? min(v, o) : max(v, o);
      }
    if (lane < n) sidx[b + lane] = v;
  } else if (n > 64) {
    for (int i = lane; i < n; i += 64) {
      const int v = sidx[b + i];
      int rank = 0;
      for (int k = 0; k < n; k++) rank += sidx[b + k] < v;
      tmp[b + rank] = v;
    }
  }
  __syncthreads();
  if (n > 64)
    for (int i = lane; i < n; i += 64) sidx[b + i] = tmp[b + i];
}

__global__ void k_bgather(int nreal, const int *__restrict__ sidx, const double *__restrict__ r,
                          const double *__restrict__ q, const int *__restrict__ type,
                          const int *__restrict__ sysid, const int *__restrict__ cellid,
                          const DSys *__restrict__ sy, double4 *__restrict__ spos,
                          int4 *__restrict__ sinfo, int *__restrict__ scell) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nreal) return;
  const int i = sidx[j], s = sysid[i];
  const DSys &S = sy[s];
  spos[j] = make_double4(wrapL(r[3 * i], S.L, S.iL), wrapL(r[3 * i + 1], S.L, S.iL),
                         wrapL(r[3 * i + 2], S.L, S.iL), q[i]);
  sinfo[j] = make_int4(type[i], i, s, 0);
  scell[j] = cellid[i];
}

// ---------------------------------------------------------------- r-space
template <typename R>
__device__ inline void bspline(const R *e, double rr, double sgrid, R &eru, R &erd, bool want_u);

template <>
__device__ inline void bspline<double>(const double *e, double rr, double sgrid, double &eru,
                                       double &erd, bool want_u) {
  const double *p = e + 8 * (int)(rr * sgrid);
  erd = p[4] / (p[5] + rr) + p[3];
  if (want_u) eru = p[1] / (p[2] + rr) + p[0];
}

template <>
__device__ inline void bspline<float>(const float *e, double rr, double sgrid, float &eru,
                                      float &erd, bool want_u) {
  const int i = (int)(rr * sgrid);
  const float *p = e + 8 * i;
  const float t = (float)(rr - i / sgrid);
  erd = p[3] - p[4] * t * __frcp_rn(p[5] + t);
  if (want_u) eru = p[0] - p[1] * t * __frcp_rn(p[2] + t);
}

template <typename R>
__device__ inline R brcp(R x);
template <>
__device__ inline float brcp<float>(float x) { return __frcp_rn(x); }
template <>
__device__ inline double brcp<double>(double x) { return 1.0 / x; }

// one lane per sorted site over the (2w+1)^3 cells of its system's grid;
// forces to the site's own row (no atomics), per-site sums halved later;
// E6: the exp-6 table [tau eu ef ru C1q C2q A A4] (sim/exp6/sitesite.h:59-66)
template <typename R, bool MEASURE, bool E6>
__global__ __launch_bounds__(kBlock) void k_bpair(
    int nreal, const DSys *__restrict__ sy, const double4 *__restrict__ spos,
    const int4 *__restrict__ sinfo, const int *__restrict__ scell, const int *__restrict__ cstart,
    const R *__restrict__ lj, const R *__restrict__ erf, double *__restrict__ f,
    double *__restrict__ epart) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= nreal) return;
  const double4 pi = spos[s];
  const int4 ii = sinfo[s];
  const DSys S = sy[ii.z];
  const int nc = S.nc, w = S.w, c = scell[s] - S.cell_off;
  const int cz = c % nc, cy = (c / nc) % nc, cx = c / (nc * nc);
  const R qi = (R)pi.w;
  const R *ljrow = lj + S.lj_off + (size_t)ii.x * S.nt * 8;
  const R *et = erf + S.erf_off;
  double fx = 0, fy = 0, fz = 0, U = 0, Eel = 0, Vir = 0, pv[6] = {0, 0, 0, 0, 0, 0};
  double npair = 0, nlj = 0;
  for (int dx = -w; dx <= w; dx++) {
    int nx = cx + dx;
    double sx = 0;
    if (nx < 0) { nx += nc; sx = -S.L; } else if (nx >= nc) { nx -= nc; sx = S.L; }
    for (int dy = -w; dy <= w; dy++) {
      int ny = cy + dy;
      double sy_ = 0;
      if (ny < 0) { ny += nc; sy_ = -S.L; } else if (ny >= nc) { ny -= nc; sy_ = S.L; }
      R ac[3] = {0, 0, 0};
      // the z cells of the column in dz order, as runs of consecutive cells
      // with one image shift (a run is a contiguous range of sorted sites);
      // per run 32 candidates at a time are tested into a bit mask and the
      // pair terms run only for the set bits, in the order of the cells
      auto wrapz = [&](int z, double &sh) {
        sh = 0;
        if (z < 0) { z += nc; sh = -S.L; } else if (z >= nc) { z -= nc; sh = S.L; }
        return z;
      };
      for (int dz = -w; dz <= w;) {
        double sz;
        const int nz = wrapz(cz + dz, sz);
        int n1 = nz;
        for (dz++; dz <= w; dz++) {
          double sh;
          const int z = wrapz(cz + dz, sh);
          if (z != n1 + 1 || sh != sz) break;
          n1 = z;
        }
        const int c0 = S.cell_off + (nx * nc + ny) * nc;
        const int jlo = cstart[c0 + nz], jhi = cstart[c0 + n1 + 1];
        for (int jb = jlo; jb < jhi; jb += 32) {
          unsigned mask = 0;
          const int nb = min(32, jhi - jb);
          for (int t = 0; t < nb; t++) {
            const double4 pj = spos[jb + t];
            const double dxr = pi.x - (pj.x + sx), dyr = pi.y - (pj.y + sy_), dzr = pi.z - (pj.z + sz);
            const double rr = dxr * dxr + dyr * dyr + dzr * dzr;
            mask |= (unsigned)(rr < S.cq && jb + t != s) << t;
          }
          while (mask) {
            const int j = jb + __ffs(mask) - 1;
            mask &= mask - 1;
            const double4 pj = spos[j];
            const double dxr = pi.x - (pj.x + sx), dyr = pi.y - (pj.y + sy_), dzr = pi.z - (pj.z + sz);
            const double rr = dxr * dxr + dyr * dyr + dzr * dzr;
            npair += 1;
            const R *ss = ljrow + sinfo[j].x * 8;
            R fr = 0, u = 0;
            if (rr < (double)ss[5]) {
              if (rr < (double)ss[4]) {
                if (E6) {
                  const R q = (R)rr, x = sqrt(q), y = exp(ss[0] * x), iz = brcp<R>(q * q * q);
                  if (MEASURE) u = ss[1] * y - ss[3] * iz;
                  fr = ss[2] * y * brcp<R>(x) - (R)6 * ss[3] * iz * brcp<R>(q);
                } else {
                  const R x = ss[0] * brcp<R>((R)rr), y = x * x * x;
                  if (MEASURE) {
                    const R z = y * y;
                    u = ss[1] * (z - y);
                    fr = ss[2] * (z + z - y) * x;
                  } else {
                    fr = ss[3] * y * (y - (R)0.5) * x;
                  }
                }
              } else {
                const R x = (R)(rr - (double)ss[5]);
                if (MEASURE) u = x * x * ss[6];
                fr = ss[7] * x;
              }
              if (MEASURE) Vir += rr * (double)fr;
              nlj += ss[6] != 0;
            }
            const R qq = qi * (R)pj.w;
            if (qq != 0) {
              R eru = 0, erd;
              bspline<R>(et, rr, S.sgrid, eru, erd, MEASURE);
              fr += qq * erd;
              if (MEASURE) Eel += (double)(qq * eru);
            }
            const R ddx = (R)dxr, ddy = (R)dyr, ddz = (R)dzr;
            const R fdx = fr * ddx, fdy = fr * ddy, fdz = fr * ddz;
            ac[0] += fdx; ac[1] += fdy; ac[2] += fdz;
            if (MEASURE) {
              U += u;
              pv[0] += (double)(ddx * fdx); pv[1] += (double)(ddy * fdy); pv[2] += (double)(ddz * fdz);
              pv[3] += (double)(ddz * fdy); pv[4] += (double)(ddx * fdz); pv[5] += (double)(ddy * fdx);
            }
          }
        }
      }
      fx += ac[0]; fy += ac[1]; fz += ac[2];
    }
  }
  f[3 * ii.y] = fx;
  f[3 * ii.y + 1] = fy;
  f[3 * ii.y + 2] = fz;
  double *e = epart + (size_t)s * kNE;
  e[0] = 0.5 * U; e[1] = 0.5 * Eel; e[2] = 0.5 * Vir;
  for (int k = 0; k < 6; k++) e[3 + k] = 0.5 * pv[k];
  e[9] = 0.5 * npair; e[10] = 0.5 * nlj;
}

// per-system sums of the per-site pair sums (sorted sites of a system are
// contiguous), one workgroup per system, fixed order
__global__ void k_bsum_sites(const DSys *__restrict__ sy, const double *__restrict__ epart,
                             double *__restrict__ sacc) {
  __shared__ double red[kBlock / 64][kNE];
  const DSys S = sy[blockIdx.x];
  double v[kNE];
#pragma unroll
  for (int k = 0; k < kNE; k++) v[k] = 0;
  for (int j = S.srt_off + threadIdx.x; j < S.srt_off + S.ns; j += kBlock)
#pragma unroll
    for (int k = 0; k < kNE; k++) v[k] += epart[(size_t)j * kNE + k];
  const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kNE; k++) {
    const double t = wave_sum(v[k]);
    if (ln == 0) red[wv][k] = t;
  }
  __syncthreads();
  if (threadIdx.x < kNE) {
    double t = 0;
    for (int w = 0; w < kBlock / 64; w++) t += red[w][threadIdx.x];
    sacc[(size_t)blockIdx.x * B_N + threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------- MD
// leap-frog of one site (sim/shakev3.c for a molecule without constraints)
// and per-wave kinetic sums; a wave holds one system
__global__ __launch_bounds__(kBlock) void k_bleap(int npad, const int *__restrict__ type,
                                                  const int *__restrict__ sysid,
                                                  const double *__restrict__ stepp,
                                                  const double *__restrict__ mass,
                                                  double *__restrict__ r, double *__restrict__ r1,
                                                  const double *__restrict__ f,
                                                  double *__restrict__ wpart) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (i < npad && type[i] >= 0) {
    const double h = stepp[sysid[i] * 5], m = mass[i], im = h * h / m;
    double V1[3], V2[3];
    for (int d = 0; d < 3; d++) {
      V1[d] = r1[3 * i + d];
      V2[d] = im * f[3 * i + d] + V1[d];
      r[3 * i + d] += V2[d];
      r1[3 * i + d] = V2[d];
    }
    v[0] = (V1[0] * V1[0] + V1[1] * V1[1] + V1[2] * V1[2] + V2[0] * V2[0] + V2[1] * V2[1] + V2[2] * V2[2]) * m;
    v[1] = (V1[0] * V1[0] + V2[0] * V2[0]) * m;
    v[2] = (V1[1] * V1[1] + V2[1] * V2[1]) * m;
    v[3] = (V1[2] * V1[2] + V2[2] * V2[2]) * m;
    v[4] = (V1[1] * V1[2] + V2[1] * V2[2]) * m;
    v[5] = (V1[2] * V1[0] + V2[2] * V2[0]) * m;
    v[6] = (V1[0] * V1[1] + V2[0] * V2[1]) * m;
  }
  const int lane = threadIdx.x & 63, wave = i >> 6;
  for (int k = 0; k < 7; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0 && wave < npad / kWS) wpart[(size_t)wave * 7 + k] = s;
  }
}

// per system: kinetic sums of its waves, then the Berendsen factor
// (sim/thermo.c:141-160 with the clamp of sim/constrd.c:98-107)
__global__ void k_bsum_kin(const DSys *__restrict__ sy, const double *__restrict__ wpart,
                           const double *__restrict__ stepp, double *__restrict__ sacc,
                           double *__restrict__ lam) {
  const DSys S = sy[blockIdx.x];
  const int w0 = S.pad_off / kWS, w1 = (S.pad_off + S.ns + kWS - 1) / kWS;
  __shared__ double red[kBlock];
  __shared__ double tot[7];
  for (int k = 0; k < 7; k++) {
    double v = 0;
    for (int w = w0 + threadIdx.x; w < w1; w += kBlock) v += wpart[(size_t)w * 7 + k];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) { sacc[(size_t)blockIdx.x * B_N + B_KIN + k] = red[0]; tot[k] = red[0]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double *p = stepp + blockIdx.x * 5;   // h, T, tau, nf, thermostat
    double l = 1;
    if (p[4] == 1) {
      const double Tk = tot[0] * 0.5 / (p[0] * p[0]) / p[3];
      const double e2 = 7.38905609893065022723;
      const double b = Tk > p[1] * e2 ? 1 : (Tk * e2 < p[1] ? -1 : log(Tk / p[1]) / 2);
      l = exp(b * -p[0] / p[2]);
    }
    lam[blockIdx.x] = l;
  }
}

__global__ void k_bscale(int npad, const int *__restrict__ sysid, const double *__restrict__ lam,
                         double *__restrict__ r1) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= npad) return;
  const double l = lam[sysid[i]];
  for (int d = 0; d < 3; d++) r1[3 * i + d] *= l;
}

template <typename R, bool E6>
void launch_bpair(Batch &b, int measure, const DSys *sy, const double4 *sp, const int4 *si,
                  const R *lj, const R *erf) {
  if (measure)
    k_bpair<R, true, E6><<<nblk(b.nreal), kBlock, 0, b.stream>>>(
        b.nreal, sy, sp, si, b.scell.p, b.cstart.p, lj, erf, b.f.p, b.epart.p);
  else
    k_bpair<R, false, E6><<<nblk(b.nreal), kBlock, 0, b.stream>>>(
        b.nreal, sy, sp, si, b.scell.p, b.cstart.p, lj, erf, b.f.p, b.epart.p);
}

// ---------------------------------------------------------------- host
int forces_launch(Batch &b, int measure) {
  const DSys *sy = (const DSys *)b.dsys.p;
  if (b.timing) b.tm.start(T_BIN, b.stream);
  k_bzero_int<<<nblk(b.ncell), kBlock, 0, b.stream>>>(b.cnt.p, b.ncell);
  k_bcell<<<nblk(b.npad), kBlock, 0, b.stream>>>(b.npad, b.r.p, b.type.p, b.sysid.p, sy, b.cellid.p,
                                                  b.cnt.p);
  k_bscan_sys<<<b.nsys, kBlock, 0, b.stream>>>(sy, b.nsys, b.nreal, b.cnt.p, b.cstart.p, b.cursor.p);
  k_bscatter<<<nblk(b.npad), kBlock, 0, b.stream>>>(b.npad, b.type.p, b.cellid.p, b.cstart.p,
                                                     b.cursor.p, b.sidx.p);
  k_bsort<<<nblk(64L * b.ncell), kBlock, 0, b.stream>>>(b.ncell, b.cstart.p, b.sidx.p, b.scell.p);
  k_bgather<<<nblk(b.nreal), kBlock, 0, b.stream>>>(b.nreal, b.sidx.p, b.r.p, b.q.p, b.type.p,
                                                     b.sysid.p, b.cellid.p, sy,
                                                     (double4 *)b.spos.p, (int4 *)b.sinfo.p,
                                                     b.scell.p);
  // padding rows of f stay zero
  k_bfill<<<nblk(3L * b.npad), kBlock, 0, b.stream>>>(3L * b.npad, b.f.p, 0.0);
  const double4 *sp = (const double4 *)b.spos.p;
  const int4 *si = (const int4 *)b.sinfo.p;
  if (b.timing) { b.tm.stop(T_BIN, b.stream); b.tm.start(T_PAIR, b.stream); }
  if (b.mode == MDX_FP64) {
    if (b.exp6) launch_bpair<double, true>(b, measure, sy, sp, si, b.lj64.p, b.erf64.p);
    else launch_bpair<double, false>(b, measure, sy, sp, si, b.lj64.p, b.erf64.p);
  } else {
    if (b.exp6) launch_bpair<float, true>(b, measure, sy, sp, si, b.lj32.p, b.erf32.p);
    else launch_bpair<float, false>(b, measure, sy, sp, si, b.lj32.p, b.erf32.p);
  }
  k_bsum_sites<<<b.nsys, kBlock, 0, b.stream>>>(sy, b.epart.p, b.sacc.p);
  if (b.timing) b.tm.stop(T_PAIR, b.stream);
  int rc;
  if ((rc = bk_forces(b, measure))) return rc;
  if (hipGetLastError() != hipSuccess) return fail(MDX_E_GPU, "batch: launch failed");
  return 0;
}

void fill_sys_out(const Batch &b, int s, const double *a, mdx_out_t *out) {
  std::memset(out, 0, sizeof(*out));
  out->npair = (long long)std::llround(a[B_NPAIR]);
  out->nlj = (long long)std::llround(a[B_NLJ]);
  const DSys &S = b.hs[s];
  const double Ek = a[B_EK] / (4 * kPi);
  const double *M = a + B_M;
  const double Edip = S.cdip / 2 * (M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
  double Pk[6];
  for (int i = 0; i < 6; i++) Pk[i] = -2 / (4 * kPi) * a[B_PK + i];
  for (int i = 0; i < 3; i++) Pk[i] += Ek + Edip;
  Pk[0] -= S.cdip * M[0] * M[0]; Pk[1] -= S.cdip * M[1] * M[1]; Pk[2] -= S.cdip * M[2] * M[2];
  Pk[3] -= S.cdip * M[1] * M[2]; Pk[4] -= S.cdip * M[2] * M[0]; Pk[5] -= S.cdip * M[0] * M[1];
  out->En_kspace = S.ks >= 0 ? Ek + Edip - S.alpha / std::sqrt(kPi) * b.sumq2[s] + b.ebg[s] : 0.0;
  out->En_el = out->En_kspace + a[B_EEL];
  out->En_pot = a[B_U];
  out->En_vir = a[B_VIR];
  for (int i = 0; i < 6; i++) out->Pvir[i] = (S.ks >= 0 ? Pk[i] : 0.0) + a[B_PV + i];
}

// cells of edge >= cutoff/v with at least 2v+1 per side (no cell is met
// twice by the stencil); boxes under 3 cutoffs take v above the asked w
int choose_cells(DSys &S, int w) {
  for (int v = w; v <= 8; v++) {
    const int nc = (int)std::floor(S.L * v / std::sqrt(S.cq));
    if (nc >= 2 * v + 1) { S.nc = nc; S.w = v; S.ncell = nc * nc * nc; return 0; }
  }
  return fail(MDX_E_ARG, "batch: box too small for the cutoff");
}

int layout_cells(Batch &b, int w) {
  int off = 0, rc;
  for (auto &S : b.hs) {
    if ((rc = choose_cells(S, w))) return rc;
    S.cell_off = off;
    off += S.ncell;
  }
  b.ncell = off;
  if ((rc = b.cnt.alloc(b.ncell)) || (rc = b.cstart.alloc(b.ncell + 1)) || (rc = b.cursor.alloc(b.ncell)) ||
      (rc = b.dsys.alloc(sizeof(DSys) * b.nsys)))
    return rc;
  MDX_HIP(hipMemcpy(b.dsys.p, b.hs.data(), sizeof(DSys) * b.nsys, hipMemcpyHostToDevice));
  b.graph_ok = false;
  return 0;
}

void release(Batch &b) {
  DBuf<double> *d[] = {&b.r, &b.f, &b.r1, &b.q, &b.mass, &b.spos, &b.lj64, &b.erf64,
                       &b.epart, &b.sacc, &b.lam, &b.stepp};
  for (auto *x : d) x->release();
  DBuf<int> *n[] = {&b.type, &b.sysid, &b.cellid, &b.cnt, &b.cstart, &b.cursor, &b.sidx, &b.scell,
                    &b.sinfo};
  for (auto *x : n) x->release();
  b.lj32.release();
  b.erf32.release();
  b.dsys.release();
  b.k.release();
  if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.tm.release();
  if (b.stream) (void)hipStreamDestroy(b.stream);
}

}  // namespace

struct mdx_batch_s {
  Batch b;
};

extern "C" int mdx_batch_create(mdx_batch_t **out, int nsys, const mdx_system_t *sys, int mode) {
  if (!out || !sys || nsys <= 0) return fail(MDX_E_ARG, "mdx_batch_create: bad argument");
  *out = nullptr;
  if (mode != MDX_FP64 && mode != MDX_MIXED) return fail(MDX_E_ARG, "mdx_batch_create: bad mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(MDX_E_NODEV, "mdx_batch_create: no HIP device");
  auto *h = new mdx_batch_s;
  Batch &b = h->b;
  b.mode = mode;
  b.k.kmfma = mode == MDX_MIXED ? 2 : 0;   // pass 2 on the split-FP16 matrix cores by default
  b.nsys = nsys;
  auto bail = [&](int rc) { release(b); delete h; return rc; };
  (void)hipGetDevice(&b.device);
  if (hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(MDX_E_GPU, "mdx_batch_create: stream"));
  // host layout
  std::vector<double> lj, erf;
  std::vector<float> lj32, erf32;
  int pad = 0, srt = 0;
  b.hs.resize(nsys);
  b.sumq2.assign(nsys, 0.0);
  b.ebg.assign(nsys, 0.0);
  b.kappa.assign(nsys, 0.0);
  for (int s = 0; s < nsys; s++) {
    const mdx_system_t &y = sys[s];
    if ((y.sstab_exp6 != nullptr) != (sys[0].sstab_exp6 != nullptr))
      return bail(fail(MDX_E_ARG, "mdx_batch_create: all systems need the same site-site form "
                                  "(system " + std::to_string(s) + " differs from system 0)"));
    if (y.ns <= 0 || y.nmol != y.ns || y.chargepol || y.ncons != 0 || !y.mass || !y.charge ||
        !y.site_type || !(y.sstab || y.sstab_exp6) || !y.erfc_tab || y.erfc_n <= 0)
      return bail(fail(MDX_E_ARG, "mdx_batch_create: systems must be atomic (one site per molecule, "
                                  "no constraints, no Drudes) with masses, charges and tables"));
    if (y.L[0] != y.L[1] || y.L[0] != y.L[2]) return bail(fail(MDX_E_ARG, "mdx_batch_create: cubic boxes only"));
    if (!(y.cutoff > 0) || y.cutoff * 2 > y.L[0]) return bail(fail(MDX_E_ARG, "mdx_batch_create: cutoff > L/2"));
    DSys &S = b.hs[s];
    std::memset(&S, 0, sizeof(S));
    S.L = y.L[0];
    S.iL = 1 / S.L;
    S.cq = y.cutoff * y.cutoff;
    S.sgrid = y.erfc_sgrid;
    S.alpha = y.alpha;
    S.cdip = 4 * kPi / (2 * y.epsinf + 1) / (S.L * S.L * S.L);
    S.nt = y.ntypes;
    S.lj_off = (int)lj.size();
    b.exp6 = y.sstab_exp6 != nullptr;
    std::vector<double> t = b.exp6 ? exp6_table(y.sstab_exp6, y.ntypes * y.ntypes)
                                   : lj_table(y.sstab, y.ntypes * y.ntypes);
    lj.insert(lj.end(), t.begin(), t.end());
    // spline intervals up to the cutoff (index (int)(r^2 sgrid), r^2 <
    // cutoff^2); none for a system without charges
    double qnet = 0;
    for (int i = 0; i < y.ns; i++) {
      b.sumq2[s] += y.charge[i] * y.charge[i];
      qnet += y.charge[i];
    }
    // Ewald corrections: the slab form is not carried by the batch; a net
    // charge adds the background energy, el_diag the diagonal term
    if (y.el_corr)
      return bail(fail(MDX_E_ARG, "mdx_batch_create: the Yeh-Berkowitz correction (el_corr) is not "
                                  "supported in a batch"));
    if (int rc0 = ewald_check(0, y.el_bg, y.epsinf, qnet, b.sumq2[s], "mdx_batch_create")) return bail(rc0);
    if (b.sumq2[s] > 0) {
      b.ebg[s] = -kPi / 2 * (qnet / y.alpha) * (qnet / y.alpha) / (S.L * S.L * S.L);
      if (y.el_diag == 1) b.ebg[s] += ewald_ediag(y.L, y.kappa, y.alpha, b.sumq2[s]);
    }
    const int nerf = b.sumq2[s] > 0 ? (int)(S.cq * y.erfc_sgrid) + 1 : 0;
    if (nerf > y.erfc_n) return bail(fail(MDX_E_ARG, "mdx_batch_create: spline table shorter than the cutoff"));
    S.erf_off = (int)erf.size();
    for (int i = 0; i < nerf; i++) {
      const mdx_erfc_t &p = y.erfc_tab[i];
      const double v[8] = {p.Au, p.Bu, p.Cu, p.Ad, p.Bd, p.Cd, 0, 0};
      erf.insert(erf.end(), v, v + 8);
      const double x0 = i / y.erfc_sgrid;
      const double Du = p.Cu + x0, bu = p.Bu / Du, Dd = p.Cd + x0, bd = p.Bd / Dd;
      const float w[8] = {(float)(bu + p.Au), (float)bu, (float)Du, (float)(bd + p.Ad), (float)bd,
                          (float)Dd, 0, 0};
      erf32.insert(erf32.end(), w, w + 8);
    }
    S.pad_off = pad;
    S.ns = y.ns;
    S.srt_off = srt;
    srt += y.ns;
    pad += (y.ns + kWS - 1) / kWS * kWS;
    S.ks = -1;
    if (b.sumq2[s] > 0 && !(y.kappa > 0))
      return bail(fail(MDX_E_ARG, "mdx_batch_create: charged system needs kappa"));
    b.kappa[s] = y.kappa;
  }
  b.npad = pad;
  lj32.assign(lj.begin(), lj.end());
  // per-site padded data
  std::vector<double> q(pad, 0.0), mass(pad, 1.0);
  std::vector<int> type(pad, -1), sysid(pad, 0);
  int nreal = 0;
  for (int s = 0; s < nsys; s++) {
    const mdx_system_t &y = sys[s];
    const int o = b.hs[s].pad_off, e = o + (y.ns + kWS - 1) / kWS * kWS;
    for (int i = o; i < e; i++) sysid[i] = s;
    for (int i = 0; i < y.ns; i++) {
      q[o + i] = y.charge[i];
      mass[o + i] = y.mass[i];
      type[o + i] = y.site_type[i];
      if (y.site_type[i] < 0 || y.site_type[i] >= y.ntypes)
        return bail(fail(MDX_E_ARG, "mdx_batch_create: site type out of range"));
    }
    nreal += y.ns;
  }
  b.nreal = nreal;
  int rc;
  if ((rc = b.r.alloc(3L * pad)) || (rc = b.f.alloc(3L * pad)) || (rc = b.r1.alloc(3L * pad)) ||
      (rc = b.q.alloc(pad)) || (rc = b.mass.alloc(pad)) || (rc = b.type.alloc(pad)) ||
      (rc = b.sysid.alloc(pad)) || (rc = b.cellid.alloc(pad)) || (rc = b.sidx.alloc(nreal)) ||
      (rc = b.scell.alloc(std::max(nreal, pad))) || (rc = b.sinfo.alloc(4L * nreal)) ||
      (rc = b.spos.alloc(4L * nreal)) || (rc = b.lj64.alloc(lj.size())) || (rc = b.lj32.alloc(lj.size())) ||
      (rc = b.erf64.alloc(erf.size())) || (rc = b.erf32.alloc(erf32.size())) ||
      (rc = b.epart.alloc((size_t)kNE * nreal)) || (rc = b.sacc.alloc((size_t)B_N * nsys)) ||
      (rc = b.lam.alloc(nsys)) || (rc = b.stepp.alloc(5L * nsys)))
    return bail(rc);
  auto up = [&](void *d, const void *src, size_t n) {
    return n == 0 || hipMemcpy(d, src, n, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up(b.q.p, q.data(), sizeof(double) * pad) || !up(b.mass.p, mass.data(), sizeof(double) * pad) ||
      !up(b.type.p, type.data(), sizeof(int) * pad) || !up(b.sysid.p, sysid.data(), sizeof(int) * pad) ||
      !up(b.lj64.p, lj.data(), sizeof(double) * lj.size()) ||
      !up(b.lj32.p, lj32.data(), sizeof(float) * lj32.size()) ||
      !up(b.erf64.p, erf.data(), sizeof(double) * erf.size()) ||
      !up(b.erf32.p, erf32.data(), sizeof(float) * erf32.size()))
    return bail(fail(MDX_E_GPU, "mdx_batch_create: upload failed"));
  // zero state and sums (kernels: the step may be captured later)
  k_bfill<<<nblk(3L * pad), kBlock, 0, b.stream>>>(3L * pad, b.r.p, 0.0);
  k_bfill<<<nblk(3L * pad), kBlock, 0, b.stream>>>(3L * pad, b.r1.p, 0.0);
  k_bfill<<<nblk((long)B_N * nsys), kBlock, 0, b.stream>>>((long)B_N * nsys, b.sacc.p, 0.0);
  if (hipStreamSynchronize(b.stream) != hipSuccess) return bail(fail(MDX_E_GPU, "mdx_batch_create: init"));
  if ((rc = layout_cells(b, 1)) || (rc = bk_setup(b, 16, 8))) return bail(rc);
  *out = h;
  return 0;
}

extern "C" int mdx_batch_destroy(mdx_batch_t *h) {
  if (!h) return 0;
  (void)hipSetDevice(h->b.device);
  if (h->b.stream) (void)hipStreamSynchronize(h->b.stream);
  release(h->b);
  delete h;
  return 0;
}

namespace {
int up_rows(Batch &b, double *dst, const double *src) {
  // concatenated [sum ns][3] -> padded rows
  std::vector<double> tmp(3L * b.npad, 0.0);
  size_t o = 0;
  for (const DSys &S : b.hs) {
    std::memcpy(tmp.data() + 3L * S.pad_off, src + o, sizeof(double) * 3 * S.ns);
    o += 3L * S.ns;
  }
  MDX_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice, b.stream));
  MDX_HIP(hipStreamSynchronize(b.stream));
  return 0;
}

int down_rows(Batch &b, double *dst, const double *src) {
  std::vector<double> tmp(3L * b.npad);
  MDX_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, b.stream));
  MDX_HIP(hipStreamSynchronize(b.stream));
  size_t o = 0;
  for (const DSys &S : b.hs) {
    std::memcpy(dst + o, tmp.data() + 3L * S.pad_off, sizeof(double) * 3 * S.ns);
    o += 3L * S.ns;
  }
  return 0;
}
}  // namespace

extern "C" int mdx_batch_forces(mdx_batch_t *h, const double *r, double *f, mdx_out_t *out,
                                int measure) {
  if (!h || !r || !f || !out) return fail(MDX_E_ARG, "mdx_batch_forces: NULL argument");
  Batch &b = h->b;
  MDX_HIP(hipSetDevice(b.device));
  int rc;
  b.state_on_device = false;
  if ((rc = up_rows(b, b.r.p, r)) || (rc = forces_launch(b, measure))) return rc;
  std::vector<double> a((size_t)B_N * b.nsys);
  MDX_HIP(hipMemcpyAsync(a.data(), b.sacc.p, sizeof(double) * a.size(), hipMemcpyDeviceToHost, b.stream));
  if ((rc = down_rows(b, f, b.f.p))) return rc;
  for (int s = 0; s < b.nsys; s++) {
    fill_sys_out(b, s, a.data() + (size_t)s * B_N, &out[s]);
    if (!measure) {
      const long long np = out[s].npair, nl = out[s].nlj;
      std::memset(&out[s], 0, sizeof(out[s]));
      out[s].npair = np;
      out[s].nlj = nl;
    }
  }
  return 0;
}

extern "C" int mdx_batch_upload_state(mdx_batch_t *h, const double *r, const double *r1) {
  if (!h || !r || !r1) return fail(MDX_E_ARG, "mdx_batch_upload_state: NULL argument");
  Batch &b = h->b;
  MDX_HIP(hipSetDevice(b.device));
  int rc;
  if ((rc = up_rows(b, b.r.p, r)) || (rc = up_rows(b, b.r1.p, r1))) return rc;
  b.state_on_device = true;
  return 0;
}

extern "C" int mdx_batch_download_state(mdx_batch_t *h, double *r, double *r1) {
  if (!h || !r || !r1) return fail(MDX_E_ARG, "mdx_batch_download_state: NULL argument");
  Batch &b = h->b;
  if (!b.state_on_device) return fail(MDX_E_ARG, "mdx_batch_download_state: no state");
  MDX_HIP(hipSetDevice(b.device));
  int rc;
  if ((rc = down_rows(b, r, b.r.p)) || (rc = down_rows(b, r1, b.r1.p))) return rc;
  return 0;
}

namespace {
int step_launch(Batch &b, int measure) {
  int rc;
  if ((rc = forces_launch(b, measure))) return rc;
  if (b.timing) b.tm.start(T_SHAKE, b.stream);
  k_bleap<<<nblk(b.npad), kBlock, 0, b.stream>>>(b.npad, b.type.p, b.sysid.p, b.stepp.p, b.mass.p,
                                                  b.r.p, b.r1.p, b.f.p, b.epart.p);
  k_bsum_kin<<<b.nsys, kBlock, 0, b.stream>>>((const DSys *)b.dsys.p, b.epart.p, b.stepp.p,
                                              b.sacc.p, b.lam.p);
  k_bscale<<<nblk(b.npad), kBlock, 0, b.stream>>>(b.npad, b.sysid.p, b.lam.p, b.r1.p);
  if (b.timing) b.tm.stop(T_SHAKE, b.stream);
  if (hipGetLastError() != hipSuccess) return fail(MDX_E_GPU, "batch: launch failed");
  return 0;
}
}  // namespace

extern "C" int mdx_batch_md_steps(mdx_batch_t *h, const mdx_step_t *p, int nsteps,
                                  mdx_step_out_t *out) {
  if (!h || !p || !out || nsteps < 0) return fail(MDX_E_ARG, "mdx_batch_md_steps: bad argument");
  Batch &b = h->b;
  if (!b.state_on_device) return fail(MDX_E_ARG, "mdx_batch_md_steps: upload the state first");
  MDX_HIP(hipSetDevice(b.device));
  std::vector<double> sp(5L * b.nsys);
  for (int s = 0; s < b.nsys; s++) {
    if (!(p[s].h > 0)) return fail(MDX_E_ARG, "mdx_batch_md_steps: h must be > 0");
    if (p[s].thermostat != 0 && p[s].thermostat != 1)
      return fail(MDX_E_ARG, "mdx_batch_md_steps: thermostat 0 (none) or 1 (Berendsen) only");
    if (p[s].thermostat && !(p[s].tau > 0 && p[s].nf > 0 && p[s].T > 0))
      return fail(MDX_E_ARG, "mdx_batch_md_steps: Berendsen needs T, tau, nf > 0");
    sp[5 * s] = p[s].h; sp[5 * s + 1] = p[s].T; sp[5 * s + 2] = p[s].tau; sp[5 * s + 3] = p[s].nf;
    sp[5 * s + 4] = p[s].thermostat;
  }
  MDX_HIP(hipMemcpyAsync(b.stepp.p, sp.data(), sizeof(double) * sp.size(), hipMemcpyHostToDevice, b.stream));
  // the per-site leap-frog reuses epart for the per-wave kinetic sums; it
  // must hold npad/64 * 7 values
  if ((size_t)b.npad / kWS * 7 > (size_t)kNE * b.nreal) return fail(MDX_E_ARG, "batch: sums buffer");
  int every = p[0].measure_every > 0 ? p[0].measure_every : nsteps;
  std::memset(out, 0, sizeof(mdx_step_out_t) * b.nsys);
  int rc = 0;
  std::vector<double> a((size_t)B_N * b.nsys), lam(b.nsys);
  for (int s = 1; s <= nsteps && !rc; s++) {
    const int measure = (s % every == 0) || s == nsteps;
    if (!measure && b.use_graph && !b.timing) {
      bool same = b.graph_ok && b.graph_p.size() == (size_t)b.nsys;
      for (int k = 0; same && k < b.nsys; k++)
        same = b.graph_p[k].h == p[k].h && b.graph_p[k].T == p[k].T && b.graph_p[k].tau == p[k].tau &&
               b.graph_p[k].nf == p[k].nf && b.graph_p[k].thermostat == p[k].thermostat;
      if (!same) {
        if (b.graph_exec) { (void)hipGraphExecDestroy(b.graph_exec); b.graph_exec = nullptr; }
        if (b.graph) { (void)hipGraphDestroy(b.graph); b.graph = nullptr; }
        b.graph_ok = false;
        MDX_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
        rc = step_launch(b, 0);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b.stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); break; }
        if (e != hipSuccess || !g) { rc = fail(MDX_E_GPU, "batch: graph capture failed"); break; }
        b.graph = g;
        if (hipGraphInstantiate(&b.graph_exec, b.graph, nullptr, nullptr, 0) != hipSuccess) {
          rc = fail(MDX_E_GPU, "batch: graph instantiation failed");
          break;
        }
        b.graph_ok = true;
        b.graph_p.assign(p, p + b.nsys);
      }
      if (hipGraphLaunch(b.graph_exec, b.stream) != hipSuccess) { rc = fail(MDX_E_GPU, "batch: graph launch"); break; }
    } else {
      rc = step_launch(b, measure);
      if (rc) break;
    }
    if (measure) {
      if (hipMemcpyAsync(a.data(), b.sacc.p, sizeof(double) * a.size(), hipMemcpyDeviceToHost, b.stream) != hipSuccess ||
          hipMemcpyAsync(lam.data(), b.lam.p, sizeof(double) * lam.size(), hipMemcpyDeviceToHost, b.stream) != hipSuccess ||
          hipStreamSynchronize(b.stream) != hipSuccess) {
        rc = fail(MDX_E_GPU, "batch: download");
        break;
      }
      for (int k = 0; k < b.nsys; k++) {
        const double *ak = a.data() + (size_t)k * B_N;
        fill_sys_out(b, k, ak, &out[k].frc);
        const double hh = p[k].h * p[k].h, hscale = 0.5 / hh;
        out[k].En_kin = ak[B_KIN] * hscale / 2;
        for (int d = 0; d < 6; d++) out[k].Pkin[d] = ak[B_PKIN + d] * hscale;
        out[k].lambda = lam[k];
      }
    }
  }
  (void)hipStreamSynchronize(b.stream);
  return rc;
}

extern "C" int mdx_batch_set_graph(mdx_batch_t *h, int on) {
  if (!h) return fail(MDX_E_ARG, "mdx_batch_set_graph: NULL");
  h->b.use_graph = on != 0;
  return 0;
}

// HIP-event timing of the kernel families on the batch stream (bin, pair,
// -, kpass1, -, kpass2, leap-frog in the slots of mdx_get_timing)
extern "C" int mdx_batch_set_timing(mdx_batch_t *h, int on) {
  if (!h) return fail(MDX_E_ARG, "mdx_batch_set_timing: NULL");
  h->b.timing = on != 0;
  return 0;
}

extern "C" int mdx_batch_get_timing(mdx_batch_t *h, double *ms, int *count, int n) {
  if (!h || !ms || !count || n < T_NFAM) return fail(MDX_E_ARG, "mdx_batch_get_timing: bad argument");
  MDX_HIP(hipStreamSynchronize(h->b.stream));
  h->b.tm.collect(ms, count);
  return 0;
}

extern "C" int mdx_batch_set_stencil(mdx_batch_t *h, int w) {
  if (!h || w < 1 || w > 4) return fail(MDX_E_ARG, "mdx_batch_set_stencil: w in 1..4");
  MDX_HIP(hipSetDevice(h->b.device));
  return layout_cells(h->b, w);
}

extern "C" int mdx_batch_set_kspace_mfma(mdx_batch_t *h, int on) {
  if (!h) return fail(MDX_E_ARG, "mdx_batch_set_kspace_mfma: NULL");
  if (on < 0 || on > 2) return fail(MDX_E_ARG, "mdx_batch_set_kspace_mfma: 0, 1 or 2");
  if (on && h->b.mode != MDX_MIXED) return fail(MDX_E_ARG, "mdx_batch_set_kspace_mfma: MDX_MIXED only");
  MDX_HIP(hipSetDevice(h->b.device));
  MDX_HIP(hipStreamSynchronize(h->b.stream));
  h->b.k.kmfma = on;
  h->b.graph_ok = false;
  return 0;
}

extern "C" int mdx_batch_set_kspace(mdx_batch_t *h, int kzc, int kparts) {
  if (!h) return fail(MDX_E_ARG, "mdx_batch_set_kspace: NULL");
  MDX_HIP(hipSetDevice(h->b.device));
  MDX_HIP(hipStreamSynchronize(h->b.stream));
  return bk_setup(h->b, kzc, kparts);
}

namespace {
// mean time of 3 force calls (after one untimed) on the device positions
int time_forces(Batch &b, hipEvent_t e0, hipEvent_t e1, double *ms) {
  int rc;
  if ((rc = forces_launch(b, 0))) return rc;
  // the faster of two timings of 3 calls each (less sensitive to a slow
  // moment of the device than one)
  *ms = 1e30;
  for (int trial = 0; trial < 2; trial++) {
    MDX_HIP(hipEventRecord(e0, b.stream));
    for (int k = 0; k < 3; k++)
      if ((rc = forces_launch(b, 0))) return rc;
    MDX_HIP(hipEventRecord(e1, b.stream));
    MDX_HIP(hipEventSynchronize(e1));
    float t = 0;
    MDX_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = std::min(*ms, (double)t / 3);
  }
  return 0;
}
}  // namespace

// the stencil w (cells >= cutoff/w, 1..3), then the k-space kz chunk
// (8..32) and parts (1..32) with the fastest force call on r;
// choice = {w, kz chunk, parts}, ms = {after the stencil, after the k-space}
extern "C" int mdx_batch_autotune(mdx_batch_t *h, const double *r, int *choice, double *ms) {
  if (!h || !r) return fail(MDX_E_ARG, "mdx_batch_autotune: NULL argument");
  Batch &b = h->b;
  MDX_HIP(hipSetDevice(b.device));
  int rc;
  b.state_on_device = false;
  if ((rc = up_rows(b, b.r.p, r))) return rc;
  hipEvent_t e0, e1;
  MDX_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(MDX_E_GPU, "mdx_batch_autotune: event");
  }
  double best = 1e30, t = 0, best_k = 1e30;
  int bw = 1, bz = b.k.kzc, bp = b.k.kparts;
  for (int w = 1; w <= 3 && !rc; w++) {
    if ((rc = layout_cells(b, w)) || (rc = time_forces(b, e0, e1, &t))) break;
    if (t < best) { best = t; bw = w; }
  }
  if (!rc) rc = layout_cells(b, bw);
  const int kzs[4] = {8, 16, 24, 32}, kps[6] = {1, 2, 4, 8, 16, 32};
  int bm = b.k.kmfma;
  // MDX_MIXED: the FP32 FMA and the two matrix-core variants of pass 2
  for (int mf = 0; mf < (b.mode == MDX_MIXED ? 3 : 1) && !rc && b.k.nks; mf++) {
    b.k.kmfma = mf;
    for (int iz = 0; iz < 4 && !rc; iz++)
      for (int ip = 0; ip < 6 && !rc; ip++) {
        if ((rc = bk_setup(b, kzs[iz], kps[ip])) || (rc = time_forces(b, e0, e1, &t))) break;
        if (t < best_k) { best_k = t; bz = kzs[iz]; bp = kps[ip]; bm = b.k.kmfma; }
      }
  }
  b.k.kmfma = bm;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = bk_setup(b, bz, bp))) return rc;
  if (choice) { choice[0] = bw; choice[1] = bz; choice[2] = bp; }
  if (ms) { ms[0] = best; ms[1] = b.k.nks ? best_k : best; }
  return 0;
}

// iv: systems, sites, padded sites, cells, charged systems, mode, w,
// charges, k segments, pass-1 units, pass-2 units, kz chunk, parts asked,
// matrix-core pass 2
extern "C" int mdx_batch_info(mdx_batch_t *h, int *iv, int n) {
  if (!h || !iv) return fail(MDX_E_ARG, "mdx_batch_info: NULL");
  const Batch &b = h->b;
  const BK &k = b.k;
  const int v[14] = {b.nsys, b.nreal, b.npad, b.ncell, k.nks, b.mode, b.hs.empty() ? 0 : b.hs[0].w,
                     k.nqt, k.nseg, k.n1, k.kmfma ? k.n2m : k.n2, k.kzc, k.kparts, k.kmfma};
  for (int i = 0; i < n && i < 14; i++) iv[i] = v[i];
  return 0;
}
