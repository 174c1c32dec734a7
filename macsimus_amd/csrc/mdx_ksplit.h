// Pass 2 of the Ewald k-space on the FP16 matrix cores with split operands,
// shared by the single-system engine (mdx_kspace.hip) and the replica
// batch (mdx_bkspace.hip): the splitting of FP32 values into FP16 parts,
// the conversion of a row group's coefficients (k_coefh) and the force
// contraction of one block of charges (kforce_h_body).  Included at file
// scope; everything here has internal linkage.
#pragma once

#include <hip/hip_runtime.h>

namespace {

constexpr int kKhBlock = 256;   // threads per workgroup of the kernels below
constexpr int kKhKB = 4;        // kz per coefficient block of the row groups (coefg)
typedef float kh_f32x16 __attribute__((ext_vector_type(16)));

__device__ inline double kh_frac1(double x) { return x - floor(x); }

// cos, sin of 2 pi t
__device__ inline void kh_cs2pi(double t, double &c, double &s) {
  sincospi(2 * kh_frac1(t), &s, &c);
}

// ------------------------------------------- pass 2, split-FP16 operands
// Each FP32 operand value x is written as the sum of NPART FP16 parts,
// x_0 = fp16(x), x_1 = fp16(x - x_0), x_2 = fp16(x - x_0 - x_1) (the
// remainders are exact in FP32); NPART = 3 holds all 24 bits of x.  The
// product x y is then the sum of x_a y_b over a + b < NPART: the terms
// left out are of relative size 2^-22 (NPART = 2) or 2^-33 (NPART = 3).
// A product of two FP16 values is exact in FP32, so the FP16 matrix
// instruction v_mfma_f32_32x32x16_f16 with FP32 accumulation forms these
// sums at 16 times the FP32 rate.  So that the parts stay in FP16's
// normal range, the coefficients of a row group are scaled by 2^(14-e),
// with e the exponent of their largest magnitude, and the phases by 2^8;
// the epilogue scales the sums back (powers of two: exact).
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

template <int NPART>
__device__ inline void split16(float x, _Float16 (&p)[NPART]) {
  float r = x;
#pragma unroll
  for (int k = 0; k < NPART; k++) {
    p[k] = (_Float16)r;
    r -= (float)p[k];
  }
}

// the coefficients of a row group (coefg, FP32, operand order of
// k_kforce_mfma) as NPART FP16 parts in 8-kz blocks: slot b of group gi,
// part k, lane l holds kz 8b..8b+7 of coefg lane l (M index l % 32, cos or
// sin by l / 32), zeros past the group's kz slots.  One workgroup per
// group; gscale[gi] = 2^(e-14) undoes the group's scale.
template <int NPART>
__global__ __launch_bounds__(kKhBlock) void k_coefh(const int *__restrict__ grp_base,
                                                  const int *__restrict__ grp_nb,
                                                  const int *__restrict__ grp_hb,
                                                  const int *__restrict__ grp_nb8,
                                                  const float *__restrict__ coefg,
                                                  half8 *__restrict__ coefh,
                                                  float *__restrict__ gscale) {
  const int gi = blockIdx.x;
  const float *src = coefg + grp_base[gi];
  const int nkz = grp_nb[gi] * kKhKB;
  float m = 0;
  for (int i = threadIdx.x; i < nkz * 64; i += kKhBlock) m = fmaxf(m, fabsf(src[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  __shared__ float red[kKhBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = red[0];
  for (int w = 1; w < kKhBlock / 64; w++) m = fmaxf(m, red[w]);
  const int e = m > 0 ? ilogbf(m) : 0;           // m in [2^e, 2^(e+1))
  const float sc = ldexpf(1.0f, 14 - e);
  if (threadIdx.x == 0) gscale[gi] = ldexpf(1.0f, e - 14);
  const int nb8 = grp_nb8[gi];
  half8 *dst = coefh + (size_t)grp_hb[gi] * NPART * 64;
  for (int t = threadIdx.x; t < nb8 * 64; t += kKhBlock) {
    const int b = t >> 6, l = t & 63;
    half8 v[NPART];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int kz = 8 * b + i;
      _Float16 p[NPART];
      split16<NPART>(kz < nkz ? src[(size_t)kz * 64 + l] * sc : 0.0f, p);
#pragma unroll
      for (int k = 0; k < NPART; k++) v[k][i] = p[k];
    }
#pragma unroll
    for (int k = 0; k < NPART; k++) dst[((size_t)b * NPART + k) * 64 + l] = v[k];
  }
}

// Pass 2 of k_kforce_mfma with split-FP16 operands.  Same work split, row
// groups, lane roles and epilogue; the K index of the instruction is (cos
// or sin by the lane half, 8 kz), so lane l supplies 8 kz of one M index's
// coefficients and 8 kz of its charge's cos (h = 0) or sin (h = 1) of c,
// and one instruction per product of parts covers 8 kz of the chunk.  The
// four waves of a workgroup take different charges through the same row
// groups, so each group's coefficient blocks are loaded once per
// workgroup into LDS (double buffered: the next group's loads are in
// flight while this group's products run; one barrier per group).
// The body is shared by the single-system kernel and the replica batch:
// charges qpos[0..nq) of one system (cubic or not), block jblk of
// 32 NTC (kKhBlock / 64) of them, work part `part` (kz chunk, kx range),
// chunk_kymax and grp_start of the system, global group tables; the
// forces of the part go to kpo[3 j].
template <int KZC, int NTC, int NPART>
__device__ inline void kforce_h_body(
    int nq, const double4 *__restrict__ qpos, double iLx, double iLy, double iLz, int kxmax,
    const int4 part, int jblk, const int *__restrict__ chunk_kymax,
    const int *__restrict__ grp_start, const int *__restrict__ grp_hb,
    const int *__restrict__ grp_nb8, const half8 *__restrict__ coefh,
    const float *__restrict__ gscale, double *__restrict__ kpo) {
  constexpr int BPC = KZC / 8;         // 8-kz blocks per chunk
  constexpr int NV = BPC * NPART * 64;  // operand vectors of a group's chunk
  constexpr int NL = (NV + kKhBlock - 1) / kKhBlock;   // of them per thread
  __shared__ half8 stage[2][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = lane & 31, h = lane >> 5;
  const int jw = (jblk * (kKhBlock / 64) + wv) * 32 * NTC + c;
  const int ch = part.x;
  const float kxs = (float)iLx, kys = (float)iLy;
  double sx[NTC], cb1[NTC], sb1[NTC], cb2[NTC], sb2[NTC], cb4[NTC], sb4[NTC];
  half8 zt[NTC][BPC][NPART];
#pragma unroll
  for (int u = 0; u < NTC; u++) {
    const double4 qp = qpos[min(jw + 32 * u, nq - 1)];
    sx[u] = kh_frac1(qp.x * iLx);
    const double sy = kh_frac1(qp.y * iLy), sz = kh_frac1(qp.z * iLz);
    kh_cs2pi(sy, cb1[u], sb1[u]);
    cb2[u] = cb1[u] * cb1[u] - sb1[u] * sb1[u];
    sb2[u] = 2 * cb1[u] * sb1[u];
    cb4[u] = cb2[u] * cb2[u] - sb2[u] * sb2[u];
    sb4[u] = 2 * cb2[u] * sb2[u];
    double cz1, sz1;
    kh_cs2pi(sz, cz1, sz1);
#pragma unroll
    for (int m = 0; m < BPC; m++) {
      double cc, ss;
      kh_cs2pi((double)(ch * KZC + 8 * m) * sz, cc, ss);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        _Float16 p[NPART];
        split16<NPART>((float)(h ? ss : cc) * 256.0f, p);
#pragma unroll
        for (int q = 0; q < NPART; q++) zt[u][m][q][k] = p[q];
        const double nc = cc * cz1 - ss * sz1;
        ss = cc * sz1 + ss * cz1;
        cc = nc;
      }
    }
  }
  const int *kym_c = chunk_kymax + ch * (kxmax + 1);
  const half8 *cfl = coefh + (size_t)ch * BPC * NPART * 64;
  // the operand vectors of group gi's chunk (blocks past the group's last
  // are not read by the products)
  auto fetch = [&](int gi, half8 (&v)[NL]) {
    const int n = min(grp_nb8[gi] - ch * BPC, BPC) * NPART * 64;
    const half8 *src = cfl + (size_t)grp_hb[gi] * NPART * 64;
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int t = threadIdx.x + i * kKhBlock;
      if (t < n) v[i] = src[t];
    }
  };
  auto put = [&](int buf, const half8 (&v)[NL]) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int t = threadIdx.x + i * kKhBlock;
      if (t < NV) stage[buf][t] = v[i];
    }
  };
  // Per kx the force of a row is kx/Lx (ca B - sa A) along x, ca C + sa D
  // along y and ca E + sa F along z, with A = cb G0 + sb G1, B = cb G2 +
  // sb G3, C = ky/Ly (cb G1 - sb G0), D = ky/Ly (cb G3 - sb G2), E = cb H0
  // + sb H1, F = cb H2 + sb H3 (ca, sa: phase of a; cb, sb: of b): A..F
  // are summed over the rows of a kx in FP32, the kx terms in FP64.
  double Fx[NTC], Fy[NTC], Fz[NTC];
  float PA[NTC], PB[NTC], PC[NTC], PD[NTC], PE[NTC], PF[NTC];
#pragma unroll
  for (int u = 0; u < NTC; u++) {
    Fx[u] = 0; Fy[u] = 0; Fz[u] = 0;
    PA[u] = 0; PB[u] = 0; PC[u] = 0; PD[u] = 0; PE[u] = 0; PF[u] = 0;
  }
  double cb[NTC], sb[NTC];
  int kx = part.y, ky0 = 0, buf = 0;
  {
    half8 v[NL];
    fetch(grp_start[kx], v);
    put(0, v);
  }
  __syncthreads();
  while (true) {
    // the group after this one: next 4 rows of kx, or the first of kx + 1
    int nkx = kx, nky0 = ky0 + 4;
    if (nky0 > kym_c[kx]) { nkx = kx + 1; nky0 = 0; }
    const bool more = nkx < part.z;
    half8 v[NL];
    if (more) fetch(grp_start[nkx] + (nky0 >> 2), v);
    if (ky0 == 0) {
#pragma unroll
      for (int u = 0; u < NTC; u++) { cb[u] = h ? cb1[u] : 1.0; sb[u] = h ? sb1[u] : 0.0; }
    }
    const float kyl = (float)(ky0 + h) * kys, kyl2 = (float)(ky0 + 2 + h) * kys;
    const int gi = grp_start[kx] + (ky0 >> 2);
    const int nb = min(grp_nb8[gi] - ch * BPC, BPC);   // wave-uniform, >= 1
    const half8 *ga = stage[buf] + lane;
    // per 8-kz block the smaller products first, then the leading one
    // (0, 0): a group's sums run over at most KZC / 8 blocks, short enough
    // that one accumulator costs no accuracy to measure (force error 4.07e-7
    // against 4.06e-7 with the leading products apart; pass 1 sums over
    // thousands of charges and keeps them apart)
    kh_f32x16 acc[NTC];
#pragma unroll
    for (int u = 0; u < NTC; u++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[u][r] = 0.0f;
#pragma unroll
    for (int b = 0; b < BPC; b++) {
      if (b < nb) {
        half8 ev[NPART];
#pragma unroll
        for (int k = 0; k < NPART; k++) ev[k] = ga[(b * NPART + k) * 64];
#pragma unroll
        for (int u = 0; u < NTC; u++) {
          if constexpr (NPART == 3) {
            acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ev[2], zt[u][b][0], acc[u], 0, 0, 0);
            acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ev[1], zt[u][b][1], acc[u], 0, 0, 0);
            acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ev[0], zt[u][b][2], acc[u], 0, 0, 0);
          }
          acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ev[1], zt[u][b][0], acc[u], 0, 0, 0);
          acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ev[0], zt[u][b][1], acc[u], 0, 0, 0);
          acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ev[0], zt[u][b][0], acc[u], 0, 0, 0);
        }
      }
    }
    // the group's scale goes into the row phases
    const float sc = gscale[gi] * (1.0f / 256.0f);
#pragma unroll
    for (int u = 0; u < NTC; u++) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const float G0 = acc[u][8 * q], G1 = acc[u][8 * q + 1], G2 = acc[u][8 * q + 2],
                    G3 = acc[u][8 * q + 3];
        const float H0 = acc[u][8 * q + 4], H1 = acc[u][8 * q + 5], H2 = acc[u][8 * q + 6],
                    H3 = acc[u][8 * q + 7];
        const double cq = q ? cb[u] * cb2[u] - sb[u] * sb2[u] : cb[u];
        const double sq = q ? cb[u] * sb2[u] + sb[u] * cb2[u] : sb[u];
        const float rcb = (float)cq * sc, rsb = (float)sq * sc;
        const float ky = q ? kyl2 : kyl;
        PA[u] = fmaf(rcb, G0, fmaf(rsb, G1, PA[u]));
        PB[u] = fmaf(rcb, G2, fmaf(rsb, G3, PB[u]));
        PC[u] = fmaf(ky, rcb * G1 - rsb * G0, PC[u]);
        PD[u] = fmaf(ky, rcb * G3 - rsb * G2, PD[u]);
        PE[u] = fmaf(rcb, H0, fmaf(rsb, H1, PE[u]));
        PF[u] = fmaf(rcb, H2, fmaf(rsb, H3, PF[u]));
      }
      const double nc = cb[u] * cb4[u] - sb[u] * sb4[u];
      sb[u] = cb[u] * sb4[u] + sb[u] * cb4[u];
      cb[u] = nc;
    }
    if (nkx != kx) {
      // the kx terms of the rows summed so far
      const float kxl = (float)kx * kxs;
#pragma unroll
      for (int u = 0; u < NTC; u++) {
        double ca, sa;
        kh_cs2pi(kx * sx[u], ca, sa);
        const float rca = (float)ca, rsa = (float)sa;
        Fx[u] += (double)(kxl * (rca * PB[u] - rsa * PA[u]));
        Fy[u] += (double)(rca * PC[u] + rsa * PD[u]);
        Fz[u] += (double)(rca * PE[u] + rsa * PF[u]);
        PA[u] = 0; PB[u] = 0; PC[u] = 0; PD[u] = 0; PE[u] = 0; PF[u] = 0;
      }
    }
    if (!more) break;
    // every wave has read stage[buf ^ 1] (last group) before the previous
    // barrier, so it can take the next group now
    put(buf ^ 1, v);
    __syncthreads();
    buf ^= 1;
    kx = nkx;
    ky0 = nky0;
  }
#pragma unroll
  for (int u = 0; u < NTC; u++) {
    const double fx = Fx[u] + __shfl_xor(Fx[u], 32, 64);
    const double fy = Fy[u] + __shfl_xor(Fy[u], 32, 64);
    const double fz = Fz[u] + __shfl_xor(Fz[u], 32, 64);
    const int j = jw + 32 * u;
    if (h || j >= nq) continue;
    double *o = kpo + (size_t)j * 3;
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
  }
}

}  // namespace
