// Ewald k-space of the batched replicas: the T formulation of the
// single-system engine (Ewald(1)/Ewald(2) of sim/ewald.c:562-1003 with
// sim/ewpass1.c, sim/ewpass2m.c:189-300) laid out per charged system and
// run for all systems by one launch per kernel.
//
// For a triple (kx,ky,kz) >= 0 and phases a, b, c of a charge the eight
// sums T_pqr = sum_j q_j P_p(a_j) P_q(b_j) P_r(c_j) (P_0 = cos, P_1 = sin)
// fix the structure factors of the four sign variants (kx, +-ky, +-kz);
// pass 1 forms them (lane <-> row (kx,ky) of a segment of 64 rows and 8 kz,
// charges in LDS chunks with cos/sin tables built by recursion), the
// coefficient kernel turns them into energy, pressure tensor and 16
// coefficients per triple, and pass 2 (lane <-> charge, kz in chunks held
// in registers, rows walked in lexicographic order) contracts them into
// the forces.  Work units carry their system: pass-1 units (system, group
// of 4 segments, charge split), pass-2 units (system, 256 charges, part of
// the kz chunks and kx range).  Partial sums over charge splits and parts
// are added in a fixed order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mdx_batch.h"
#include "mdx_ksplit.h"

using namespace mdx;
using namespace mdx::batch;

namespace {

constexpr int kBlock = 256;
constexpr int kJC = 32;          // charges per LDS chunk in pass 1
constexpr int kKB = 4;           // pass-2 coefficient rows are padded to kKB kz
constexpr int kJMin = 1024;      // charges per pass-1 split (at least)
constexpr double kPi = 3.14159265358979323846;

template <typename R>
struct C2 { R x, y; };

inline int nblk(long n, int b = kBlock) { return (int)((n + b - 1) / b); }

__device__ inline double frac1(double x) { return x - floor(x); }

__device__ inline void cs2pi(double t, double &c, double &s) { sincospi(2 * frac1(t), &s, &c); }

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// charge positions (unwrapped) and charges of the batch charge list
__global__ void k_bk_gather(int nqt, const int *__restrict__ qrow, const double *__restrict__ r,
                            const double *__restrict__ q, double4 *__restrict__ qpos) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nqt) return;
  const int i = qrow[g];
  qpos[g] = make_double4(r[3 * i], r[3 * i + 1], r[3 * i + 2], q[i]);
}

// ------------------------------------------------------------------ pass 1
// unit = (system, group of 4 segments, charge split); lane <-> row
template <typename R>
__global__ __launch_bounds__(kBlock) void k_bk_sfac(
    const int4 *__restrict__ units, const KSys *__restrict__ ks, const double4 *__restrict__ qpos,
    const int *__restrict__ seg_blk, const int *__restrict__ seg_row0,
    const int *__restrict__ seg_rowend, const int *__restrict__ row_kx,
    const int *__restrict__ row_ky, double *__restrict__ kpart, double *__restrict__ tsum) {
  extern __shared__ char smem[];
  const int4 u = units[blockIdx.x];
  const KSys K = ks[u.x];
  const int KXT = K.KXT, KYT = K.KYT, KZT = K.KZT;
  C2<R> *Xs = (C2<R> *)smem;            // [kJC][KXT], q folded in
  C2<R> *Ys = Xs + kJC * KXT;           // [kJC][KYT]
  C2<R> *Zs = Ys + kJC * KYT;           // [kJC][KZT]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int seg = u.y * 4 + wv;
  const bool segok = seg < K.nseg;
  int blk = 0, kx = 0, ky = 0;
  if (segok) {
    const int gs = K.seg_off + seg;
    blk = seg_blk[gs];
    const int row = seg_row0[gs] + lane;
    if (row < seg_rowend[gs]) { kx = row_kx[row]; ky = row_ky[row]; }
  }
  blk = __builtin_amdgcn_readfirstlane(blk);
  const int j0 = K.q_off + u.z * K.jper, j1 = min(K.q_off + K.nq, j0 + K.jper);
  const int ntx = (KXT + 7) / 8, nty = (KYT + 7) / 8, ntz = (KZT + 7) / 8, ntask = ntx + nty + ntz;
  const double iL = K.iL;

  R acc[8][8];
  double mid[8][8];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int m = 0; m < 8; m++) { acc[k][m] = 0; mid[k][m] = 0; }

  for (int jc = j0; jc < j1; jc += kJC) {
    // cos/sin tables of this chunk: anchors e^{2 pi i 8m s} and the base
    // e^{2 pi i s} in FP64, 7 recursion steps in R
    for (int t = threadIdx.x; t < kJC * ntask; t += kBlock) {
      const int jj = t / ntask, task = t % ntask;
      int d, m, KT;
      C2<R> *dst;
      if (task < ntx) { d = 0; m = task; KT = KXT; dst = Xs + jj * KXT; }
      else if (task < ntx + nty) { d = 1; m = task - ntx; KT = KYT; dst = Ys + jj * KYT; }
      else { d = 2; m = task - ntx - nty; KT = KZT; dst = Zs + jj * KZT; }
      const int j = jc + jj;
      if (j < j1) {
        const double4 qp = qpos[j];
        const double x = d == 0 ? qp.x : (d == 1 ? qp.y : qp.z);
        const double sd = frac1(x * iL);
        const double qf = d == 0 ? qp.w : 1.0;
        double c0, s0, c1, s1;
        cs2pi(8.0 * m * sd, c0, s0);
        cs2pi(sd, c1, s1);
        c0 *= qf; s0 *= qf;
        R ec = (R)c0, es = (R)s0;
        const R bc = (R)c1, bs = (R)s1;
        for (int k = 8 * m; k < min(8 * m + 8, KT); k++) {
          dst[k].x = ec;
          dst[k].y = es;
          const R nc = ec * bc - es * bs;
          es = ec * bs + es * bc;
          ec = nc;
        }
      } else {
        for (int k = 8 * m; k < min(8 * m + 8, KT); k++) { dst[k].x = 0; dst[k].y = 0; }
      }
    }
    __syncthreads();
    const C2<R> *zrow = Zs + 8 * blk;
#pragma unroll 1
    for (int jj = 0; jj < kJC; jj++) {
      const C2<R> X = Xs[jj * KXT + kx], Y = Ys[jj * KYT + ky];
      const R A[4] = {X.x * Y.x, X.x * Y.y, X.y * Y.x, X.y * Y.y};
      const C2<R> *z = zrow + jj * KZT;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const C2<R> zk = z[k];
#pragma unroll
        for (int p = 0; p < 4; p++) {
          acc[k][2 * p] += A[p] * zk.x;
          acc[k][2 * p + 1] += A[p] * zk.y;
        }
      }
    }
    __syncthreads();
    if (sizeof(R) == 4) {
#pragma unroll
      for (int k = 0; k < 8; k++)
#pragma unroll
        for (int m = 0; m < 8; m++) { mid[k][m] += (double)acc[k][m]; acc[k][m] = 0; }
    }
  }
  if (!segok) return;
  // one split: the T sums themselves
  double *out = K.nsplit == 1 ? tsum + (((size_t)K.seg_off + seg) * 64 + lane) * 64
                              : kpart + ((K.kp_off + (size_t)u.z * K.nseg + seg) * 64 + lane) * 64;
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int m = 0; m < 8; m++) out[k * 8 + m] = (double)mid[k][m] + (double)acc[k][m];
}

// tsum of a segment = sum of its system's charge splits, in split order
// (systems of one split had pass 1 write tsum)
__global__ void k_bk_ksum(long n, const KSys *__restrict__ ks, const int *__restrict__ seg_ks,
                          const double *__restrict__ kpart, double *__restrict__ tsum) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int gs = (int)(i >> 12), e = (int)(i & 4095);
  const KSys &K = ks[seg_ks[gs]];
  if (K.nsplit == 1) return;
  const int ls = gs - K.seg_off;
  double s = 0;
  for (int p = 0; p < K.nsplit; p++) s += kpart[((size_t)K.kp_off + (size_t)p * K.nseg + ls) * 4096 + e];
  tsum[i] = s;
}

// ------------------------------------------------- energy, tensor, coefs
// one thread per (segment, lane, kz in block), 512 per segment, so a
// workgroup holds one system; energy and tensor (sim/ewpass2m.c:194-222
// written with the T sums) to kcpart[workgroup][7]
template <typename R>
__global__ __launch_bounds__(kBlock) void k_bk_kcoef(
    int nseg, const KSys *__restrict__ ks, const int *__restrict__ seg_ks,
    const int *__restrict__ seg_blk, const int *__restrict__ seg_row0,
    const int *__restrict__ seg_rowend, const int *__restrict__ row_kx,
    const int *__restrict__ row_ky, const int *__restrict__ row_kzmax,
    const int *__restrict__ row_coef0, const double *__restrict__ tsum, R *__restrict__ coef,
    double *__restrict__ kcpart, int measure, const int *__restrict__ grp_start,
    const int *__restrict__ grp_base, float *__restrict__ coefg) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int seg = t >> 9, lane = (t >> 3) & 63, k = t & 7;
  double E = 0, pk[6] = {0, 0, 0, 0, 0, 0};
  if (seg < nseg) {
    const KSys &K = ks[seg_ks[seg]];
    const int row = seg_row0[seg] + lane;
    const int kz = 8 * seg_blk[seg] + k;
    if (row < seg_rowend[seg] && kz <= row_kzmax[row]) {
      const int kx = row_kx[row], ky = row_ky[row];
      R *cf = coef + ((size_t)row_coef0[row] + kz) * 16;
      if (kx == 0 && ky == 0 && kz == 0) {
        for (int m = 0; m < 16; m++) cf[m] = 0;
      } else {
        const double *T = tsum + ((size_t)seg * 64 + lane) * 64 + k * 8;
        const int nz = (kx > 0) + (ky > 0) + (kz > 0);
        const double w = (double)(1 << (nz - 1)) * 0.25;
        const double kxl = kx * K.iL, kyl = ky * K.iL, kzl = kz * K.iL;
        const double kk = kxl * kxl + kyl * kyl + kzl * kzl;
        const double PIaq = -(kPi / K.alpha) * (kPi / K.alpha);
        const double e = 4 / (kk * (K.L * K.L * K.L)) * exp(PIaq * kk);
        const double we4 = 4 * w * e;
        for (int m = 0; m < 4; m++) {
          cf[2 * m] = (R)(we4 * T[2 * m]);
          cf[2 * m + 1] = (R)(we4 * T[2 * m + 1]);
          cf[8 + 2 * m] = (R)(kzl * we4 * T[2 * m + 1]);
          cf[8 + 2 * m + 1] = (R)(-kzl * we4 * T[2 * m]);
        }
        if (coefg) {
          // the same 16 values in the operand order of the matrix-core pass
          // 2 (row q = ky % 4 of group ky / 4, output p, cos/sin kk -> lane
          // 16(q/2) + 8(p/4) + 4(q%2) + p%4 + 32 kk of kz slot kz)
          const int q = ky & 3;
          float *g = coefg + grp_base[grp_start[K.grps_off + kx] + (ky >> 2)] + (size_t)kz * 64;
          const int ib = 16 * (q >> 1) + 4 * (q & 1);
          for (int e2 = 0; e2 < 16; e2++) {
            const int pp = e2 >> 1, kk2 = e2 & 1;
            g[ib + 8 * (pp >> 2) + (pp & 3) + 32 * kk2] = (float)cf[e2];
          }
        }
        if (measure) {
          double s2 = 0;
          for (int m = 0; m < 8; m++) s2 += T[m] * T[m];
          const double ptf = 1 / kk - PIaq;
          const double E1 = w * e * 4 * s2;
          double sst = 0, st = 0, ss = 0;
          for (int p = 0; p < 2; p++)
            sst += T[p * 4 + 3] * T[p * 4 + 0] - T[p * 4 + 2] * T[p * 4 + 1];
          for (int q = 0; q < 2; q++)
            st += T[4 + q * 2 + 1] * T[q * 2 + 0] - T[4 + q * 2 + 0] * T[q * 2 + 1];
          for (int rr = 0; rr < 2; rr++) ss += T[6 + rr] * T[rr] - T[4 + rr] * T[2 + rr];
          E = E1;
          pk[0] = E1 * ptf * kxl * kxl;
          pk[1] = E1 * ptf * kyl * kyl;
          pk[2] = E1 * ptf * kzl * kzl;
          pk[3] = w * e * ptf * (-8 * sst) * kyl * kzl;
          pk[4] = w * e * ptf * (-8 * st) * kxl * kzl;
          pk[5] = w * e * ptf * (-8 * ss) * kxl * kyl;
        }
      }
    }
  }
  if (measure) {
    __shared__ double red[kBlock / 64][7];
    const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double v[7] = {E, pk[0], pk[1], pk[2], pk[3], pk[4], pk[5]};
    for (int m = 0; m < 7; m++) {
      const double s = wave_sum(v[m]);
      if (ln == 0) red[wv][m] = s;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double s = 0;
      for (int w2 = 0; w2 < kBlock / 64; w2++) s += red[w2][threadIdx.x];
      kcpart[(size_t)blockIdx.x * 7 + threadIdx.x] = s;
    }
  }
}

// per batch system: k-space sums of its coefficient workgroups (measure)
// and its dipole M = sum q r over its charges; zeros without charges
__global__ void k_bk_sums(const DSys *__restrict__ sy, const KSys *__restrict__ ks,
                          const double *__restrict__ kcpart, const double4 *__restrict__ qpos,
                          int measure, double *__restrict__ sacc) {
  __shared__ double red[kBlock];
  const int ksi = sy[blockIdx.x].ks;
  for (int a = 0; a < 10; a++) {
    if (a < 7 && !measure) continue;
    double v = 0;
    if (ksi >= 0) {
      const KSys &K = ks[ksi];
      if (a < 7) {
        for (int b = K.kcb0 + threadIdx.x; b < K.kcb1; b += kBlock) v += kcpart[(size_t)b * 7 + a];
      } else {
        for (int j = K.q_off + threadIdx.x; j < K.q_off + K.nq; j += kBlock) {
          const double4 p = qpos[j];
          v += p.w * (a == 7 ? p.x : (a == 8 ? p.y : p.z));
        }
      }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) sacc[(size_t)blockIdx.x * B_N + (a < 7 ? B_EK + a : B_M + a - 7)] = red[0];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ pass 2
// unit = (system, block of 256 of its charges, part); lane <-> charge; the
// kz chunk's cos/sin of c live in registers, rows (kx,ky) reaching the
// chunk are walked in lexicographic order, the 16 coefficients of a triple
// are wave-uniform (scalar loads).  Partial forces per part.
// Units are listed system by system, charge block by charge block, so
// that the units in flight share the coefficients of few systems.
template <typename R, int KZC>
__global__ __launch_bounds__(kBlock) void k_bk_kforce(
    const int4 *__restrict__ units, const KSys *__restrict__ ks, const double4 *__restrict__ qpos,
    const int4 *__restrict__ parts, const int *__restrict__ chunk_kymax,
    const int *__restrict__ lex_row0, const int2 *__restrict__ lex_meta,
    const R *__restrict__ coef, double *__restrict__ kpf) {
  constexpr int BPC = KZC / kKB;
  const int4 u = units[blockIdx.x];
  const KSys K = ks[u.x];
  const int4 part = parts[u.z];          // chunk, kx begin, kx end, part of the system
  const int ch = part.x;
  const int j = u.y * kBlock + threadIdx.x;
  const double4 qp = qpos[K.q_off + min(j, K.nq - 1)];
  const double sx = frac1(qp.x * K.iL), sy = frac1(qp.y * K.iL), sz = frac1(qp.z * K.iL);
  const R kls = (R)K.iL;
  double cb1, sb1;
  cs2pi(sy, cb1, sb1);
  R zc[KZC], zs[KZC];
  {
    double cz1, sz1;
    cs2pi(sz, cz1, sz1);
#pragma unroll
    for (int m = 0; m < KZC / 8; m++) {
      double c, s;
      cs2pi((double)(ch * KZC + 8 * m) * sz, c, s);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        zc[8 * m + k] = (R)c;
        zs[8 * m + k] = (R)s;
        const double nc = c * cz1 - s * sz1;
        s = c * sz1 + s * cz1;
        c = nc;
      }
    }
  }
  const int *kym_c = chunk_kymax + K.chunk_off + ch * (K.kxmax + 1);
  const int *lr0 = lex_row0 + K.lexr_off;
  double Fx = 0, Fy = 0, Fz = 0;
  for (int kx = part.y; kx < part.z; kx++) {
    const int kym = kym_c[kx];
    double ca, sa;
    cs2pi(kx * sx, ca, sa);
    const R rca = (R)ca, rsa = (R)sa;
    const R kxl = (R)kx * kls;
    const int rbase = lr0[kx];
    double cb = 1, sb = 0;
    for (int ky = 0; ky <= kym; ky++) {
      const int2 mt = lex_meta[rbase + ky];
      const int nb = __builtin_amdgcn_readfirstlane(min(mt.x - ch * BPC, BPC));
      const R *cf = coef + ((size_t)__builtin_amdgcn_readfirstlane(mt.y) + ch * KZC) * 16;
      R G0 = 0, G1 = 0, G2 = 0, G3 = 0, H0 = 0, H1 = 0, H2 = 0, H3 = 0;
#pragma unroll
      for (int b = 0; b < BPC; b++) {
        if (b < nb) {
          const R *e = cf + b * (16 * kKB);
#pragma unroll
          for (int k = 0; k < kKB; k++) {
            const R c = zc[b * kKB + k], s = zs[b * kKB + k];
            const R *ek = e + 16 * k;
            G0 = fma(ek[0], c, G0); G0 = fma(ek[1], s, G0);
            G1 = fma(ek[2], c, G1); G1 = fma(ek[3], s, G1);
            G2 = fma(ek[4], c, G2); G2 = fma(ek[5], s, G2);
            G3 = fma(ek[6], c, G3); G3 = fma(ek[7], s, G3);
            H0 = fma(ek[8], c, H0); H0 = fma(ek[9], s, H0);
            H1 = fma(ek[10], c, H1); H1 = fma(ek[11], s, H1);
            H2 = fma(ek[12], c, H2); H2 = fma(ek[13], s, H2);
            H3 = fma(ek[14], c, H3); H3 = fma(ek[15], s, H3);
          }
        }
      }
      const R rcb = (R)cb, rsb = (R)sb;
      const R kyl = (R)ky * kls;
      Fx += (double)(kxl * (rca * (rcb * G2 + rsb * G3) - rsa * (rcb * G0 + rsb * G1)));
      Fy += (double)(kyl * (rca * (rcb * G1 - rsb * G0) + rsa * (rcb * G3 - rsb * G2)));
      Fz += (double)(rca * (rcb * H0 + rsb * H1) + rsa * (rcb * H2 + rsb * H3));
      const double nc = cb * cb1 - sb * sb1;
      sb = cb * sb1 + sb * cb1;
      cb = nc;
    }
  }
  if (j >= K.nq) return;
  double *o = kpf + ((size_t)K.kpf_off + (size_t)part.w * K.nq + j) * 3;
  o[0] = Fx;
  o[1] = Fy;
  o[2] = Fz;
}

// Pass 2 on the matrix cores (MDX_MIXED): the kernel k_kforce_mfma of the
// single-system engine with the unit's system descriptor; unit = (system,
// block of 128 of its charges, part), 32 charges per wave.  See that
// kernel for the instruction layout and the row-group operand order.
typedef float mfma_f32x16 __attribute__((ext_vector_type(16)));

template <int KZC>
__global__ __launch_bounds__(kBlock, 4) void k_bk_kforce_mfma(
    const int4 *__restrict__ units, const KSys *__restrict__ ks, const double4 *__restrict__ qpos,
    const int4 *__restrict__ parts, const int *__restrict__ chunk_kymax,
    const int *__restrict__ grp_start, const int *__restrict__ grp_base,
    const int *__restrict__ grp_nb, const float *__restrict__ coefg, double *__restrict__ kpf) {
  constexpr int BPC = KZC / kKB;
  const int4 u = units[blockIdx.x];
  const KSys K = ks[u.x];
  const int4 part = parts[u.z];
  const int ch = part.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = lane & 31, h = lane >> 5;
  const int j = (u.y * (kBlock / 64) + wv) * 32 + c;
  const double4 qp = qpos[K.q_off + min(j, K.nq - 1)];
  const double sx = frac1(qp.x * K.iL), sy = frac1(qp.y * K.iL), sz = frac1(qp.z * K.iL);
  const float kls = (float)K.iL;
  double cb1, sb1;
  cs2pi(sy, cb1, sb1);
  const double cb2 = cb1 * cb1 - sb1 * sb1, sb2 = 2 * cb1 * sb1;
  const double cb4 = cb2 * cb2 - sb2 * sb2, sb4 = 2 * cb2 * sb2;
  float zt[KZC];
  {
    double cz1, sz1;
    cs2pi(sz, cz1, sz1);
#pragma unroll
    for (int m = 0; m < KZC / 8; m++) {
      double cc, ss;
      cs2pi((double)(ch * KZC + 8 * m) * sz, cc, ss);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        zt[8 * m + k] = (float)(h ? ss : cc);
        const double nc = cc * cz1 - ss * sz1;
        ss = cc * sz1 + ss * cz1;
        cc = nc;
      }
    }
  }
  const int *kym_c = chunk_kymax + K.chunk_off + ch * (K.kxmax + 1);
  const int *gs = grp_start + K.grps_off;
  const float *cfl = coefg + lane + (size_t)ch * KZC * 64;
  double Fx = 0, Fy = 0, Fz = 0;
  for (int kx = part.y; kx < part.z; kx++) {
    const int kym = kym_c[kx];
    double ca, sa;
    cs2pi(kx * sx, ca, sa);
    const float rca = (float)ca, rsa = (float)sa;
    const float kxl = (float)kx * kls;
    const int g0 = gs[kx];
    double cb = h ? cb1 : 1.0, sb = h ? sb1 : 0.0;
    for (int ky0 = 0; ky0 <= kym; ky0 += 4) {
      const int gi = g0 + (ky0 >> 2);
      const int nb = min(grp_nb[gi] - ch * BPC, BPC);
      const float *ga = cfl + grp_base[gi];
      mfma_f32x16 acc;
#pragma unroll
      for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
      for (int b = 0; b < BPC; b++) {
        if (b < nb) {
          float av[kKB];
#pragma unroll
          for (int k = 0; k < kKB; k++) av[k] = ga[(b * kKB + k) * 64];
#pragma unroll
          for (int k = 0; k < kKB; k++)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[k], zt[b * kKB + k], acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const float G0 = acc[8 * q], G1 = acc[8 * q + 1], G2 = acc[8 * q + 2], G3 = acc[8 * q + 3];
        const float H0 = acc[8 * q + 4], H1 = acc[8 * q + 5], H2 = acc[8 * q + 6], H3 = acc[8 * q + 7];
        const double cq = q ? cb * cb2 - sb * sb2 : cb, sq = q ? cb * sb2 + sb * cb2 : sb;
        const float rcb = (float)cq, rsb = (float)sq;
        const float kyl = (float)(ky0 + 2 * q + h) * kls;
        Fx += (double)(kxl * (rca * (rcb * G2 + rsb * G3) - rsa * (rcb * G0 + rsb * G1)));
        Fy += (double)(kyl * (rca * (rcb * G1 - rsb * G0) + rsa * (rcb * G3 - rsb * G2)));
        Fz += (double)(rca * (rcb * H0 + rsb * H1) + rsa * (rcb * H2 + rsb * H3));
      }
      const double nc = cb * cb4 - sb * sb4;
      sb = cb * sb4 + sb * cb4;
      cb = nc;
    }
  }
  Fx += __shfl_xor(Fx, 32, 64);
  Fy += __shfl_xor(Fy, 32, 64);
  Fz += __shfl_xor(Fz, 32, 64);
  if (h || j >= K.nq) return;
  double *o = kpf + ((size_t)K.kpf_off + (size_t)part.w * K.nq + j) * 3;
  o[0] = Fx;
  o[1] = Fy;
  o[2] = Fz;
}

// Pass 2 with split-FP16 operands (kmfma 2): kforce_h_body of mdx_ksplit.h
// with the unit's system descriptor, one tile of 32 charges per wave;
// unit = (system, block of 128 of its charges, part)
template <int KZC, int NPART>
__global__ __launch_bounds__(kBlock, 3) void k_bk_kforce_h(
    const int4 *__restrict__ units, const KSys *__restrict__ ks, const double4 *__restrict__ qpos,
    const int4 *__restrict__ parts, const int *__restrict__ chunk_kymax,
    const int *__restrict__ grp_start, const int *__restrict__ grp_hb,
    const int *__restrict__ grp_nb8, const half8 *__restrict__ coefh,
    const float *__restrict__ gscale, double *__restrict__ kpf) {
  const int4 u = units[blockIdx.x];
  const KSys &K = ks[u.x];
  const int4 part = parts[u.z];
  kforce_h_body<KZC, 1, NPART>(K.nq, qpos + K.q_off, K.iL, K.iL, K.iL, K.kxmax, part, u.y,
                               chunk_kymax + K.chunk_off, grp_start + K.grps_off, grp_hb,
                               grp_nb8, coefh, gscale,
                               kpf + ((size_t)K.kpf_off + (size_t)part.w * K.nq) * 3);
}

// sums a charge's parts in order and adds -q grad Psi and the dipole field
__global__ void k_bk_kforce_sum(int nqt, const KSys *__restrict__ ks, const int *__restrict__ qks,
                                const int *__restrict__ qrow, const double4 *__restrict__ qpos,
                                const double *__restrict__ kpf, const double *__restrict__ sacc,
                                double *__restrict__ f) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nqt) return;
  const KSys &K = ks[qks[g]];
  const int j = g - K.q_off;
  double F[3] = {0, 0, 0};
  for (int p = 0; p < K.nkpart; p++) {
    const double *o = kpf + ((size_t)K.kpf_off + (size_t)p * K.nq + j) * 3;
    F[0] += o[0];
    F[1] += o[1];
    F[2] += o[2];
  }
  const double q = qpos[g].w;
  const double *M = sacc + (size_t)K.sys * B_N + B_M;
  const int i = qrow[g];
  for (int d = 0; d < 3; d++) f[3 * i + d] += -q * F[d] - q * K.cdip * M[d];
}

template <typename R>
int run(Batch &b, int measure) {
  BK &k = b.k;
  const KSys *ks = (const KSys *)k.ks.p;
  const DSys *sy = (const DSys *)b.dsys.p;
  const double4 *qpos = (const double4 *)k.qpos.p;
  if (b.timing) b.tm.start(T_KPASS1, b.stream);
  if (k.nks) {
    k_bk_gather<<<nblk(k.nqt), kBlock, 0, b.stream>>>(k.nqt, k.qrow.p, b.r.p, b.q.p,
                                                       (double4 *)k.qpos.p);
    const size_t shmem = sizeof(C2<R>) * kJC * (size_t)k.shmem_max;
    k_bk_sfac<R><<<k.n1, kBlock, shmem, b.stream>>>((const int4 *)k.units1.p, ks, qpos, k.seg_blk.p,
                                                    k.seg_row0.p, k.seg_rowend.p, k.row_kx.p,
                                                    k.row_ky.p, k.kpart.p, k.tsum.p);
  }
  if (b.timing) { b.tm.stop(T_KPASS1, b.stream); b.tm.start(T_KCOEF, b.stream); }
  R *coef = nullptr;
  if constexpr (sizeof(R) == 4) coef = k.coef32.p; else coef = (R *)k.coef64.p;
  if (k.nks) {
    const long nt = (long)k.nseg * 4096;
    k_bk_ksum<<<nblk(nt), kBlock, 0, b.stream>>>(nt, ks, k.seg_ks.p, k.kpart.p, k.tsum.p);
    float *coefg = nullptr;
    if constexpr (sizeof(R) == 4) coefg = k.kmfma ? k.coefg.p : nullptr;
    k_bk_kcoef<R><<<k.nkcb, kBlock, 0, b.stream>>>(
        k.nseg, ks, k.seg_ks.p, k.seg_blk.p, k.seg_row0.p, k.seg_rowend.p, k.row_kx.p, k.row_ky.p,
        k.row_kzmax.p, k.row_coef0.p, k.tsum.p, coef, k.kcpart.p, measure, k.grp_start.p,
        k.grp_base.p, coefg);
  }
  k_bk_sums<<<b.nsys, kBlock, 0, b.stream>>>(sy, ks, k.kcpart.p, qpos, measure, b.sacc.p);
  if (b.timing) { b.tm.stop(T_KCOEF, b.stream); b.tm.start(T_KPASS2, b.stream); }
  bool done2 = false;
  if constexpr (sizeof(R) == 4) {
    if (k.nks && k.kmfma == 2) {
      // the groups' coefficients as 3 FP16 parts, then the contraction
      k_coefh<3><<<k.ngrp, kBlock, 0, b.stream>>>(k.grp_base.p, k.grp_nb.p, k.grp_hb.p, k.grp_nb8.p,
                                                k.coefg.p, (half8 *)k.coefh.p, k.gscale.p);
#define MDX_BKH(KZV)                                                                          \
  k_bk_kforce_h<KZV, 3><<<k.n2m, kBlock, 0, b.stream>>>(                                     \
      (const int4 *)k.units2m.p, ks, qpos, (const int4 *)k.parts.p, k.chunk_kymax.p,          \
      k.grp_start.p, k.grp_hb.p, k.grp_nb8.p, (const half8 *)k.coefh.p, k.gscale.p, k.kpf.p)
      switch (k.kzc) {
        case 8: MDX_BKH(8); break;
        case 16: MDX_BKH(16); break;
        case 24: MDX_BKH(24); break;
        case 32: MDX_BKH(32); break;
        default: return fail(MDX_E_ARG, "batch k-space: kz chunk must be 8, 16, 24 or 32");
      }
#undef MDX_BKH
      k_bk_kforce_sum<<<nblk(k.nqt), kBlock, 0, b.stream>>>(k.nqt, ks, k.qks.p, k.qrow.p, qpos,
                                                             k.kpf.p, b.sacc.p, b.f.p);
      done2 = true;
    } else if (k.nks && k.kmfma) {
#define MDX_BKM(KZV)                                                                          \
  k_bk_kforce_mfma<KZV><<<k.n2m, kBlock, 0, b.stream>>>((const int4 *)k.units2m.p, ks, qpos,  \
                                                       (const int4 *)k.parts.p,              \
                                                       k.chunk_kymax.p, k.grp_start.p,       \
                                                       k.grp_base.p, k.grp_nb.p, k.coefg.p,  \
                                                       k.kpf.p)
      switch (k.kzc) {
        case 8: MDX_BKM(8); break;
        case 16: MDX_BKM(16); break;
        case 24: MDX_BKM(24); break;
        case 32: MDX_BKM(32); break;
        default: return fail(MDX_E_ARG, "batch k-space: kz chunk must be 8, 16, 24 or 32");
      }
#undef MDX_BKM
      k_bk_kforce_sum<<<nblk(k.nqt), kBlock, 0, b.stream>>>(k.nqt, ks, k.qks.p, k.qrow.p, qpos,
                                                             k.kpf.p, b.sacc.p, b.f.p);
      done2 = true;
    }
  }
  if (k.nks && !done2) {
#define MDX_BKF(KZV)                                                                          \
  k_bk_kforce<R, KZV><<<k.n2, kBlock, 0, b.stream>>>((const int4 *)k.units2.p, ks, qpos,      \
                                                     (const int4 *)k.parts.p, k.chunk_kymax.p, \
                                                     k.lex_row0.p, (const int2 *)k.lex_meta.p, \
                                                     coef, k.kpf.p)
    switch (k.kzc) {
      case 8: MDX_BKF(8); break;
      case 16: MDX_BKF(16); break;
      case 24: MDX_BKF(24); break;
      case 32: MDX_BKF(32); break;
      default: return fail(MDX_E_ARG, "batch k-space: kz chunk must be 8, 16, 24 or 32");
    }
#undef MDX_BKF
    k_bk_kforce_sum<<<nblk(k.nqt), kBlock, 0, b.stream>>>(k.nqt, ks, k.qks.p, k.qrow.p, qpos,
                                                           k.kpf.p, b.sacc.p, b.f.p);
  }
  if (b.timing) b.tm.stop(T_KPASS2, b.stream);
  MDX_HIP(hipGetLastError());
  return 0;
}

}  // namespace

namespace mdx {
namespace batch {

void BK::release() {
  ks.release();
  DBuf<int> *n[] = {&qrow, &qks, &row_kx, &row_ky, &row_kzmax, &row_coef0, &seg_blk, &seg_row0,
                    &seg_rowend, &seg_ks, &units1, &units2, &units2m, &parts, &lex_row0, &lex_meta,
                    &chunk_kymax, &grp_start, &grp_base, &grp_nb, &grp_hb, &grp_nb8};
  for (auto *x : n) x->release();
  DBuf<double> *d[] = {&qpos, &kpart, &tsum, &kpf, &kcpart, &coef64};
  for (auto *x : d) x->release();
  coef32.release();
  coefg.release();
  coefh.release();
  gscale.release();
}

int bk_setup(Batch &b, int kzc, int kparts) {
  if (kzc != 8 && kzc != 16 && kzc != 24 && kzc != 32)
    return fail(MDX_E_ARG, "batch k-space: kz chunk must be 8, 16, 24 or 32");
  if (kparts < 1 || kparts > 256) return fail(MDX_E_ARG, "batch k-space: parts in 1..256");
  BK &k = b.k;
  std::vector<KSys> hks;
  std::vector<int> qrow, qks, row_kx, row_ky, row_kzmax, row_coef0, seg_blk, seg_row0, seg_rowend,
      seg_ks, units1, units2, units2m, parts, lex_row0, lex_meta, chunk_kymax, grp_start, grp_base,
      grp_nb, grp_hb, grp_nb8;
  long grp_floats = 0, grp_h8 = 0;
  std::vector<double> cost1;
  std::vector<KLayout> lay;
  // the k-sets first: the largest sets the parts of the others
  for (int s = 0; s < b.nsys; s++) {
    b.hs[s].ks = -1;
    if (b.sumq2[s] > 0) {
      lay.emplace_back();
      const double L3[3] = {b.hs[s].L, b.hs[s].L, b.hs[s].L};
      klayout_build(L3, b.kappa[s], lay.back());
    }
  }
  long ntrip_max = 1;
  for (const KLayout &K : lay) ntrip_max = std::max<long>(ntrip_max, K.ntrip);
  int nrows = 0, nseg = 0, maxlds = 0;
  long ncoef = 0, kp_off = 0, kpf_off = 0;
  std::vector<double> qcharge(b.npad);
  // charges of the padded rows (host copy)
  if (hipMemcpy(qcharge.data(), b.q.p, sizeof(double) * b.npad, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(MDX_E_GPU, "batch k-space: charge download");
  int li = 0;
  for (int s = 0; s < b.nsys; s++) {
    if (b.sumq2[s] <= 0) continue;
    const KLayout &K = lay[li++];
    DSys &S = b.hs[s];
    KSys x{};
    x.L = S.L;
    x.iL = S.iL;
    x.alpha = S.alpha;
    x.cdip = S.cdip;
    x.sys = s;
    x.q_off = (int)qrow.size();
    for (int i = S.pad_off; i < S.pad_off + S.ns; i++)
      if (qcharge[i] != 0) { qrow.push_back(i); qks.push_back((int)hks.size()); }
    x.nq = (int)qrow.size() - x.q_off;
    x.kxmax = K.kxmax;
    x.KXT = K.kxmax + 1;
    x.KYT = *std::max_element(K.row_ky.begin(), K.row_ky.end()) + 1;
    x.KZT = K.kzpad;
    maxlds = std::max(maxlds, x.KXT + x.KYT + x.KZT);
    x.nseg = K.nseg;
    x.seg_off = nseg;
    x.nsplit = std::max(1, std::min(64, (x.nq + kJMin - 1) / kJMin));
    x.jper = (x.nq + x.nsplit - 1) / x.nsplit;
    x.kp_off = kp_off;
    if (x.nsplit > 1) kp_off += (long)x.nsplit * x.nseg;
    for (int r = 0; r < K.nrows; r++) {
      row_kx.push_back(K.row_kx[r]);
      row_ky.push_back(K.row_ky[r]);
      row_kzmax.push_back(K.row_kzmax[r]);
      row_coef0.push_back((int)(K.row_coef0[r] + ncoef));
    }
    for (int g = 0; g < K.nseg; g++) {
      seg_blk.push_back(K.seg_blk[g]);
      seg_row0.push_back(K.seg_row0[g] + nrows);
      seg_rowend.push_back(K.seg_rowend[g] + nrows);
      seg_ks.push_back((int)hks.size());
    }
    for (int g = 0; g < (K.nseg + 3) / 4; g++)
      for (int p = 0; p < x.nsplit; p++) {
        units1.insert(units1.end(), {(int)hks.size(), g, p, 0});
        cost1.push_back((double)x.jper);
      }
    // pass 2: lexicographic rows and chunk tables
    x.lexr_off = (int)lex_row0.size();
    for (int kx = 0; kx <= K.kxmax; kx++) lex_row0.push_back(K.lex_row0[kx] + nrows);
    for (int r = 0; r < K.nrows; r++) {
      lex_meta.push_back(K.lex_meta[2 * r]);
      lex_meta.push_back((int)(K.lex_meta[2 * r + 1] + ncoef));
    }
    const int np_want = std::max(1, (int)std::lround((double)kparts * K.ntrip / ntrip_max));
    std::vector<int> kym, pt;
    int nchunk = 0, rc;
    if ((rc = kchunks_build(K, kzc, np_want, kym, pt, nchunk))) return rc;
    x.chunk_off = (int)chunk_kymax.size();
    chunk_kymax.insert(chunk_kymax.end(), kym.begin(), kym.end());
    x.nkpart = (int)pt.size() / 4;
    x.kpf_off = kpf_off;
    kpf_off += (long)x.nkpart * x.nq;
    const int p0 = (int)parts.size() / 4;
    for (int p = 0; p < x.nkpart; p++)
      parts.insert(parts.end(), {pt[4 * p], pt[4 * p + 1], pt[4 * p + 2], p});
    for (int cb = 0; cb < (x.nq + kBlock - 1) / kBlock; cb++)
      for (int p = 0; p < x.nkpart; p++) {
        units2.insert(units2.end(), {(int)hks.size(), cb, p0 + p, 0});
      }
    // matrix-core pass 2: units of 128 charges; row groups of the system
    for (int cb = 0; cb < (x.nq + 127) / 128; cb++)
      for (int p = 0; p < x.nkpart; p++) units2m.insert(units2m.end(), {(int)hks.size(), cb, p0 + p, 0});
    x.grps_off = (int)grp_start.size();
    const int g0 = (int)grp_base.size();
    for (int v : K.grp_start) grp_start.push_back(v + g0);
    for (int v : K.grp_base) grp_base.push_back((int)(v + grp_floats));
    grp_nb.insert(grp_nb.end(), K.grp_nb.begin(), K.grp_nb.end());
    grp_floats += K.grp_floats;
    for (int v : K.grp_hb) grp_hb.push_back((int)(v + grp_h8));
    grp_nb8.insert(grp_nb8.end(), K.grp_nb8.begin(), K.grp_nb8.end());
    grp_h8 += K.grp_h8;
    x.kcb0 = 2 * nseg;
    x.kcb1 = 2 * (nseg + K.nseg);
    nrows += K.nrows;
    nseg += K.nseg;
    ncoef += K.ncoef;
    S.ks = (int)hks.size();
    hks.push_back(x);
  }
  // pass-1 units longest first (stable: ties keep the system order), so
  // that the short units of small systems fill the end of the launch
  auto by_cost = [](std::vector<int> &u, const std::vector<double> &c) {
    std::vector<int> ord(c.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return c[a] > c[b]; });
    std::vector<int> out;
    out.reserve(u.size());
    for (int i : ord) out.insert(out.end(), u.begin() + 4 * i, u.begin() + 4 * i + 4);
    u.swap(out);
  };
  by_cost(units1, cost1);
  k.release();
  k.hks = hks;
  k.nks = (int)hks.size();
  k.nqt = (int)qrow.size();
  k.nseg = nseg;
  k.n1 = (int)units1.size() / 4;
  k.n2 = (int)units2.size() / 4;
  k.n2m = (int)units2m.size() / 4;
  k.ngrp = (int)grp_hb.size();
  k.nkcb = 2 * nseg;
  k.kzc = kzc;
  k.kparts = kparts;
  k.shmem_max = maxlds;
  k.ncoef = ncoef;
  const size_t shmem64 = 16 * (size_t)kJC * maxlds;
  if (shmem64 > 160 * 1024) return fail(MDX_E_ARG, "batch k-space: k-set too large for the pass-1 LDS tables");
  int rc;
  auto one = [](size_t n) { return n ? n : 1; };
  if ((rc = k.ks.alloc(one(sizeof(KSys) * hks.size()))) || (rc = k.qrow.alloc(one(qrow.size()))) ||
      (rc = k.qks.alloc(one(qks.size()))) || (rc = k.qpos.alloc(one(4 * qrow.size()))) ||
      (rc = k.row_kx.alloc(one(row_kx.size()))) || (rc = k.row_ky.alloc(one(row_ky.size()))) ||
      (rc = k.row_kzmax.alloc(one(row_kzmax.size()))) ||
      (rc = k.row_coef0.alloc(one(row_coef0.size()))) || (rc = k.seg_blk.alloc(one(seg_blk.size()))) ||
      (rc = k.seg_row0.alloc(one(seg_row0.size()))) ||
      (rc = k.seg_rowend.alloc(one(seg_rowend.size()))) || (rc = k.seg_ks.alloc(one(seg_ks.size()))) ||
      (rc = k.units1.alloc(one(units1.size()))) || (rc = k.units2.alloc(one(units2.size()))) ||
      (rc = k.units2m.alloc(one(units2m.size()))) || (rc = k.grp_start.alloc(one(grp_start.size()))) ||
      (rc = k.grp_base.alloc(one(grp_base.size()))) || (rc = k.grp_nb.alloc(one(grp_nb.size()))) ||
      (rc = k.parts.alloc(one(parts.size()))) || (rc = k.lex_row0.alloc(one(lex_row0.size()))) ||
      (rc = k.lex_meta.alloc(one(lex_meta.size()))) ||
      (rc = k.chunk_kymax.alloc(one(chunk_kymax.size()))) ||
      (rc = k.kpart.alloc(one((size_t)kp_off * 4096))) || (rc = k.tsum.alloc(one((size_t)nseg * 4096))) ||
      (rc = k.kpf.alloc(one((size_t)kpf_off * 3))) || (rc = k.kcpart.alloc(one((size_t)k.nkcb * 7))))
    return rc;
  if (b.mode == MDX_MIXED) {
    if ((rc = k.coef32.alloc(one((size_t)ncoef * 16))) || (rc = k.coefg.alloc(one((size_t)grp_floats))) ||
        (rc = k.coefh.alloc(one((size_t)grp_h8 * 3 * 512))) ||
        (rc = k.gscale.alloc(one(grp_hb.size()))) || (rc = k.grp_hb.alloc(one(grp_hb.size()))) ||
        (rc = k.grp_nb8.alloc(one(grp_nb8.size()))))
      return rc;
    if (grp_h8 * 3 * 512 >= (1L << 31)) return fail(MDX_E_ARG, "batch k-space: too many row groups");
    MDX_HIP(hipMemset(k.coef32.p, 0, sizeof(float) * one((size_t)ncoef * 16)));
    MDX_HIP(hipMemset(k.coefg.p, 0, sizeof(float) * one((size_t)grp_floats)));
  } else {
    if ((rc = k.coef64.alloc(one((size_t)ncoef * 16)))) return rc;
    MDX_HIP(hipMemset(k.coef64.p, 0, sizeof(double) * one((size_t)ncoef * 16)));
  }
  MDX_HIP(hipMemset(k.kcpart.p, 0, sizeof(double) * one((size_t)k.nkcb * 7)));
  auto up = [](void *d, const void *src, size_t n) {
    return n == 0 || hipMemcpy(d, src, n, hipMemcpyHostToDevice) == hipSuccess;
  };
  auto upi = [&](DBuf<int> &d, const std::vector<int> &v) { return up(d.p, v.data(), sizeof(int) * v.size()); };
  if (!up(k.ks.p, hks.data(), sizeof(KSys) * hks.size()) || !upi(k.qrow, qrow) || !upi(k.qks, qks) ||
      !upi(k.row_kx, row_kx) || !upi(k.row_ky, row_ky) || !upi(k.row_kzmax, row_kzmax) ||
      !upi(k.row_coef0, row_coef0) || !upi(k.seg_blk, seg_blk) || !upi(k.seg_row0, seg_row0) ||
      !upi(k.seg_rowend, seg_rowend) || !upi(k.seg_ks, seg_ks) || !upi(k.units1, units1) ||
      !upi(k.units2, units2) || !upi(k.units2m, units2m) || !upi(k.grp_start, grp_start) ||
      !upi(k.grp_base, grp_base) || !upi(k.grp_nb, grp_nb) ||
      (b.mode == MDX_MIXED && (!upi(k.grp_hb, grp_hb) || !upi(k.grp_nb8, grp_nb8))) ||
      !upi(k.parts, parts) || !upi(k.lex_row0, lex_row0) ||
      !upi(k.lex_meta, lex_meta) || !upi(k.chunk_kymax, chunk_kymax))
    return fail(MDX_E_GPU, "batch k-space: upload failed");
  // the systems' k-space indices
  MDX_HIP(hipMemcpy(b.dsys.p, b.hs.data(), sizeof(DSys) * b.nsys, hipMemcpyHostToDevice));
  b.graph_ok = false;
  return 0;
}

int bk_forces(Batch &b, int measure) {
  return b.mode == MDX_FP64 ? run<double>(b, measure) : run<float>(b, measure);
}

}  // namespace batch
}  // namespace mdx
