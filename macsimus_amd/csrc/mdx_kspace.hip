// Ewald k-space on MI355X: structure-factor sums, energy / pressure tensor,
// and the per-charge forces.
//
// Reference: Ewald(1) and Ewald(2) of sim/ewald.c:562-1003 with
// sim/ewpass1.c (Q(k) sums by recursive phase products over the half
// space kx>=0 with the four sign variants of (ky,kz)), sim/ewpass2m.c:
// 189-300 (energy, analytic pressure tensor, forces) and sim/ewaldf.c.
//
// Formulation used here.  For a triple (kx,ky,kz) of non-negative integers
// with phases a=2pi kx x/Lx, b=2pi ky y/Ly, c=2pi kz z/Lz of a charge, the
// four structure factors Q(kx,+-ky,+-kz) of the reference are fixed by the
// eight real sums
//     T_pqr = sum_j q_j P_p(a_j) P_q(b_j) P_r(c_j),   P_0=cos, P_1=sin,
// (e.g. Re Q(kx,s ky,t kz) = T_ccc - st T_css - s T_ssc - t T_scs), and
//     sum_{s,t} |Q|^2 = 4 sum T_pqr^2,
//     force on j = -q_j grad_j sum_triples w e(k)/(2pi) 4 sum T_pqr P_p P_q P_r,
// where e(k) = 4 exp(-pi^2 k^2/alpha^2)/(k^2 V) is the reference's ekk and w
// = 2^(#nonzero components)/8 removes the sign variants that coincide for
// zero components.  Pass 1 is then a real GEMM: rows (kx,ky) x 4 cos/sin
// products of a and b, columns kz x {cos,sin} of c, summed over charges.
// Pass 2 contracts per charge the coefficients 4 w e T over kz and takes
// the gradient analytically.  In MDX_MIXED both run in FP32 with FP64
// accumulation across charge chunks, each as FMA chains or on the matrix
// cores (k_sfac_mfma, k_kforce_mfma); FP64 throughout in MDX_FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "mdx_internal.h"
#include "mdx_ksplit.h"

namespace {

constexpr int kBlock = 256;
constexpr int kJC = 32;          // charges per LDS chunk in pass 1
constexpr int kKB = 4;           // pass-2 coefficient rows are padded to kKB kz
constexpr double kPi = 3.14159265358979323846;

template <typename R>
struct C2 { R x, y; };

inline int nblk(long n, int b = kBlock) { return (int)((n + b - 1) / b); }

__device__ inline double frac1(double x) { return x - floor(x); }

// cos, sin of 2 pi t in FP64
__device__ inline void cs2pi(double t, double &c, double &s) { sincospi(2 * frac1(t), &s, &c); }

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// positions and charges of the charge list: index i < ns is the centre of
// site i, ns+i its Drude at r_i + rpol_i (POLAR); with them, per direction
// d the phases e^{2 pi i s} and e^{2 pi i 8 s} of s = frac(x_d / L_d) in
// FP64 (qtrig[j][d] = cos s, sin s, cos 8s, sin 8s), from which pass 1
// builds its tables
__global__ void k_qgather(int nq, int ns, const int *__restrict__ qsite,
                          const double *__restrict__ r, const double *__restrict__ charge,
                          const double *__restrict__ chargepol, double iLx, double iLy,
                          double iLz, double4 *__restrict__ qpos, double4 *__restrict__ qtrig,
                          float2 *__restrict__ qanc, float2 *__restrict__ qbase, int3 nb3,
                          int3 boff3) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nq) return;
  const int k = qsite[j];
  double4 p;
  if (k < ns) {
    p = make_double4(r[3 * k], r[3 * k + 1], r[3 * k + 2], charge[k]);
  } else {
    const int i = k - ns;
    p = make_double4(r[3 * i] + r[3 * k], r[3 * i + 1] + r[3 * k + 1],
                     r[3 * i + 2] + r[3 * k + 2], chargepol[i]);
  }
  qpos[j] = p;
  const double x[3] = {p.x * iLx, p.y * iLy, p.z * iLz};
  const int nb[3] = {nb3.x, nb3.y, nb3.z}, boff[3] = {boff3.x, boff3.y, boff3.z};
  for (int d = 0; d < 3; d++) {
    const double sd = frac1(x[d]);
    double c1, s1, c8, s8;
    cs2pi(sd, c1, s1);
    cs2pi(8.0 * sd, c8, s8);
    qtrig[3 * (size_t)j + d] = make_double4(c1, s1, c8, s8);
    if (!qanc) continue;
    // the FP32 table anchors of the matrix-core pass 1: e^{2 pi i 8m s}
    // (times q along x) for the nb[d] blocks of 8, by the same FP64
    // products as the tables of k_sfac, and the FP32 base e^{2 pi i s};
    // stored [task][charge]
    const double qf = d == 0 ? p.w : 1.0;
    double c0 = 1, s0 = 0;
    for (int m = 0; m < nb[d]; m++) {
      qanc[(size_t)(boff[d] + m) * nq + j] = make_float2((float)(c0 * qf), (float)(s0 * qf));
      const double nc = c0 * c8 - s0 * s8;
      s0 = c0 * s8 + s0 * c8;
      c0 = nc;
    }
    qbase[(size_t)d * nq + j] = make_float2((float)c1, (float)s1);
  }
}

// ------------------------------------------------------------------ pass 1
// blockIdx.x: group of 4 segments (a segment = 64 rows of one kz block);
// blockIdx.y: charge split.  Lane <-> row, the charge loop is shared, the
// kz block of a wave is uniform so its cos/sin of c are broadcast reads.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_sfac(
    int nq, const double4 *__restrict__ qpos, double iLx, double iLy, double iLz, int nseg,
    const int *__restrict__ seg_blk, const int *__restrict__ seg_row0,
    const int *__restrict__ seg_rowend, const int *__restrict__ row_kx,
    const int *__restrict__ row_ky, int KXT, int KYT, int KZT, int jper,
    double *__restrict__ kpart, const double4 *__restrict__ qtrig) {
  extern __shared__ char smem[];
  C2<R> *Xs = (C2<R> *)smem;            // [kJC][KXT], q folded in
  C2<R> *Ys = Xs + kJC * KXT;           // [kJC][KYT]
  C2<R> *Zs = Ys + kJC * KYT;           // [kJC][KZT]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int seg = blockIdx.x * 4 + wv;
  const bool segok = seg < nseg;
  int blk = 0, kx = 0, ky = 0;
  if (segok) {
    blk = seg_blk[seg];
    const int row = seg_row0[seg] + lane;
    if (row < seg_rowend[seg]) { kx = row_kx[row]; ky = row_ky[row]; }
  }
  blk = __builtin_amdgcn_readfirstlane(blk);
  const int j0 = blockIdx.y * jper, j1 = min(nq, j0 + jper);
  const int ntx = (KXT + 7) / 8, nty = (KYT + 7) / 8, ntz = (KZT + 7) / 8, ntask = ntx + nty + ntz;

  // R=float: the sums of one chunk (32 charges) in acc are added to the
  // FP64 mid after every chunk.  R=double: acc only.
  R acc[8][8];
  double mid[8][8];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int m = 0; m < 8; m++) { acc[k][m] = 0; mid[k][m] = 0; }

  for (int jc = j0; jc < j1; jc += kJC) {
    // cos/sin tables of this chunk: anchors e^{2 pi i 8m s} = (e^{2 pi i
    // 8s})^m in FP64 from the gathered phases, 7 recursion steps in R with
    // the base e^{2 pi i s}
    for (int t = threadIdx.x; t < kJC * ntask; t += kBlock) {
      const int jj = t / ntask, task = t % ntask;
      int d, m, KT;
      C2<R> *dst;
      if (task < ntx) { d = 0; m = task; KT = KXT; dst = Xs + jj * KXT; }
      else if (task < ntx + nty) { d = 1; m = task - ntx; KT = KYT; dst = Ys + jj * KYT; }
      else { d = 2; m = task - ntx - nty; KT = KZT; dst = Zs + jj * KZT; }
      const int j = jc + jj;
      if (j < j1) {
        const double4 tr = qtrig[3 * (size_t)j + d];
        const double qf = d == 0 ? qpos[j].w : 1.0;
        double c0 = 1, s0 = 0;
        for (int i = 0; i < m; i++) {
          const double nc = c0 * tr.z - s0 * tr.w;
          s0 = c0 * tr.w + s0 * tr.z;
          c0 = nc;
        }
        const double c1 = tr.x, s1 = tr.y;
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
  double *out = kpart + (((size_t)blockIdx.y * nseg + seg) * 64 + lane) * 64;
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int m = 0; m < 8; m++) out[k * 8 + m] = (double)mid[k][m] + (double)acc[k][m];
}

// Pass 1 on the matrix cores (mixed mode): the same sums as k_sfac<float>
// as a chain of v_mfma_f32_16x16x4_f32 per chunk, with the same LDS tables,
// FP64 flush per chunk and kpart layout.  A segment (64 rows x one 8-wide
// kz block) is taken by 64/(4 NT) waves of NT tiles of 4 rows each (NT = 16:
// one wave, 4 segments per workgroup; NT = 8: two waves, fewer registers
// and 2 segments per workgroup); a tile's product
// is D[i][n] = sum_k A[i][k] B[k][n] over 4 charges k, with M index i =
// (row 4t + i/4, product pq = i%4 of cos/sin a and b) and N index n = (kz
// 8 blk + n/2, cos/sin c = n%2).  Operand layout of the instruction: lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and receives D[4(l/16) + r][l%16]
// in register r.
typedef float mfma_f32x4 __attribute__((ext_vector_type(4)));

template <int NT>
__global__ __launch_bounds__(kBlock) void k_sfac_mfma(
    int nq, const double4 *__restrict__ qpos, int nseg, const int *__restrict__ seg_blk,
    const int *__restrict__ seg_row0, const int *__restrict__ seg_rowend,
    const int *__restrict__ row_kx, const int *__restrict__ row_ky, int KXT, int KYT, int KZT,
    int jper, double *__restrict__ kpart, const double4 *__restrict__ qtrig) {
  extern __shared__ char smem[];
  C2<float> *Xs = (C2<float> *)smem;    // [kJC][KXT], q folded in
  C2<float> *Ys = Xs + kJC * KXT;       // [kJC][KYT]
  C2<float> *Zs = Ys + kJC * KYT;       // [kJC][KZT]
  const float *Xf = (const float *)Xs, *Yf = (const float *)Ys, *Zf = (const float *)Zs;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  constexpr int WPS = 16 / NT;          // waves per segment
  const int seg = blockIdx.x * (4 / WPS) + wv / WPS;
  const int rbase = (wv % WPS) * 4 * NT;  // first row of this wave in the segment
  const bool segok = seg < nseg;
  int blk = 0, row0 = 0, rowend = 0;
  if (segok) { blk = seg_blk[seg]; row0 = seg_row0[seg]; rowend = seg_rowend[seg]; }
  blk = __builtin_amdgcn_readfirstlane(blk);
  // A operand of tile t: row row0 + rbase + 4t + i16/4, X factor cos/sin by
  // bit 1 of i16, Y factor by bit 0; float offsets within a charge's table row
  int xoff[NT], yoff[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int row = row0 + rbase + 4 * t + (i16 >> 2);
    int kx = 0, ky = 0;
    if (row < rowend) { kx = row_kx[row]; ky = row_ky[row]; }
    xoff[t] = 2 * kx + ((i16 >> 1) & 1);
    yoff[t] = 2 * ky + (i16 & 1);
  }
  const int zoff = 2 * (8 * blk + (i16 >> 1)) + (i16 & 1);
  const int j0 = blockIdx.y * jper, j1 = min(nq, j0 + jper);
  const int ntx = (KXT + 7) / 8, nty = (KYT + 7) / 8, ntz = (KZT + 7) / 8, ntask = ntx + nty + ntz;

  mfma_f32x4 acc[NT];
  double mid[NT][4];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) { acc[t][r] = 0.0f; mid[t][r] = 0; }

  for (int jc = j0; jc < j1; jc += kJC) {
    // the tables of k_sfac (zero for charges past the split)
    for (int t = threadIdx.x; t < kJC * ntask; t += kBlock) {
      const int jj = t / ntask, task = t % ntask;
      int d, m, KT;
      C2<float> *dst;
      if (task < ntx) { d = 0; m = task; KT = KXT; dst = Xs + jj * KXT; }
      else if (task < ntx + nty) { d = 1; m = task - ntx; KT = KYT; dst = Ys + jj * KYT; }
      else { d = 2; m = task - ntx - nty; KT = KZT; dst = Zs + jj * KZT; }
      const int j = jc + jj;
      if (j < j1) {
        const double4 tr = qtrig[3 * (size_t)j + d];
        const double qf = d == 0 ? qpos[j].w : 1.0;
        double c0 = 1, s0 = 0;
        for (int i = 0; i < m; i++) {
          const double nc = c0 * tr.z - s0 * tr.w;
          s0 = c0 * tr.w + s0 * tr.z;
          c0 = nc;
        }
        c0 *= qf; s0 *= qf;
        float ec = (float)c0, es = (float)s0;
        const float bc = (float)tr.x, bs = (float)tr.y;
        for (int k = 8 * m; k < min(8 * m + 8, KT); k++) {
          dst[k].x = ec;
          dst[k].y = es;
          const float nc = ec * bc - es * bs;
          es = ec * bs + es * bc;
          ec = nc;
        }
      } else {
        for (int k = 8 * m; k < min(8 * m + 8, KT); k++) { dst[k].x = 0; dst[k].y = 0; }
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < kJC; s += 4) {
      const int jj = s + kq;
      const float b = Zf[jj * 2 * KZT + zoff];
      const float *xr = Xf + jj * 2 * KXT, *yr = Yf + jj * 2 * KYT;
#pragma unroll
      for (int t = 0; t < NT; t++)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[xoff[t]] * yr[yoff[t]], b, acc[t], 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) { mid[t][r] += (double)acc[t][r]; acc[t][r] = 0.0f; }
  }
  if (!segok) return;
  // lane l holds rows rbase + 4t + l/16, kz 8 blk + i16/2, T index 2r + i16%2
  double *out = kpart + ((size_t)blockIdx.y * nseg + seg) * 64 * 64 + (i16 >> 1) * 8 + (i16 & 1);
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) out[(rbase + 4 * t + kq) * 64 + 2 * r] = mid[t][r];
}

// Pass 1 as a GEMM on the matrix cores (mixed mode), T[(row,pq),(kz,r)] =
// sum_j A[(row,pq),j] B[j,(kz,r)] with A = q X_p(kx) Y_q(ky), B = Z_r(kz):
// a wave owns a unit of 8 consecutive rows of the kz-sorted row list (two
// M tiles of 4 rows x 4 products) and every 8-kz block its first (longest)
// row reaches (NA <= 8 N tiles of 8 kz x cos/sin), so each A operand is
// used by all N tiles and each B operand by both M tiles; one
// v_mfma_f32_16x16x4_f32 takes 4 charges.  Layout of the instruction:
// lane l supplies A[l%16][l/16] and B[l/16][l%16] and receives D[4(l/16) +
// r][l%16] in register r, i.e. row l/16 of the tile, kz l%16/2, cos/sin
// l%2, product pq = r: T index 2r + l%2 of that row and kz.
// The cos/sin tables of a chunk of JC charges are built in LDS as in
// k_sfac (FP64 anchors every 8 entries, FP32 recursion), laid out [k][j]
// (charge fastest) with a row stride of 2 JC + 8 floats: the lanes of a
// wave build consecutive charges, so the table writes are conflict-free,
// and the B operands of a tile (8 kz x 2 x 4 charges) fall in 64 distinct
// banks.  The matrix accumulators are added to FP64 sums after each chunk.
// Output: the kpart layout of k_sfac (segments of 64 rows of one 8-kz
// block; rows of a unit past a block's rows are dropped).
template <int JC>
struct SfTab {
  static constexpr int S = 2 * JC + 8;   // floats per table row (one k)
};

// the cos/sin tables of charges [jc, jc+JC) of the split ending at j1,
// from the anchors of k_qgather: 7 FP32 recursion steps per block of 8
template <int JC, int BLK>
__device__ inline void sf_tables(float *Xf, float *Yf, float *Zf, int KXT, int KYT, int KZT,
                                 int jc, int j1, int nq, const float2 *__restrict__ qanc,
                                 const float2 *__restrict__ qbase) {
  constexpr int S = SfTab<JC>::S;
  const int ntx = (KXT + 7) / 8, nty = (KYT + 7) / 8, ntz = (KZT + 7) / 8, ntask = ntx + nty + ntz;
  for (int t = threadIdx.x; t < JC * ntask; t += BLK) {
    const int jj = t % JC, task = t / JC;
    int d, m, KT;
    float *dst;
    if (task < ntx) { d = 0; m = task; KT = KXT; dst = Xf; }
    else if (task < ntx + nty) { d = 1; m = task - ntx; KT = KYT; dst = Yf; }
    else { d = 2; m = task - ntx - nty; KT = KZT; dst = Zf; }
    dst += 2 * jj;
    const int j = jc + jj;
    const int k1 = min(8 * m + 8, KT);
    if (j < j1) {
      const float2 a = qanc[(size_t)task * nq + j], b = qbase[(size_t)d * nq + j];
      float ec = a.x, es = a.y;
      for (int k = 8 * m; k < k1; k++) {
        *(float2 *)(dst + k * S) = make_float2(ec, es);
        const float nc = ec * b.x - es * b.y;
        es = ec * b.y + es * b.x;
        ec = nc;
      }
    } else {
      for (int k = 8 * m; k < k1; k++) *(float2 *)(dst + k * S) = make_float2(0.0f, 0.0f);
    }
  }
}

// one wave's NA N tiles of a unit, the tiles t0, t0 + ts, ... (ts waves
// share a unit), over the charges [j0, j1): every wave of the workgroup
// (BLK threads) builds tables and meets the barriers alike
template <int NA, int JC, int BLK>
__device__ inline void sf_unit(float *Xf, float *Yf, float *Zf, int KXT, int KYT, int KZT,
                               int j0, int j1, int r0, int nrows, const int *__restrict__ row_kx,
                               const int *__restrict__ row_ky, const int *__restrict__ blk_rows,
                               const int *__restrict__ blk_seg0, int nseg,
                               double *__restrict__ kpart, int nq,
                               const float2 *__restrict__ qanc,
                               const float2 *__restrict__ qbase, int t0, int ts) {
  constexpr int S = SfTab<JC>::S;
  const int lane = threadIdx.x & 63, i16 = lane & 15, kq = lane >> 4;
  // operand offsets (floats) for charge 0 of a chunk's first K step:
  // A of M tile m: row r0 + 4m + i16/4, X factor by bit 1 of i16, Y by bit 0;
  // B of N tile t: kz 8t + i16/2, cos/sin by bit 0
  int xo[2], yo[2];
#pragma unroll
  for (int m = 0; m < 2; m++) {
    const int row = r0 + 4 * m + (i16 >> 2);
    int kx = 0, ky = 0;
    if (row < nrows) { kx = row_kx[row]; ky = row_ky[row]; }
    xo[m] = kx * S + 2 * kq + ((i16 >> 1) & 1);
    yo[m] = ky * S + 2 * kq + (i16 & 1);
  }
  const int zo = (i16 >> 1) * S + 2 * kq + (i16 & 1) + 8 * t0 * S;
  mfma_f32x4 acc[2][NA > 0 ? NA : 1];
  float mid[2][NA > 0 ? NA : 1][4];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int t = 0; t < NA; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) { acc[m][t][r] = 0.0f; mid[m][t][r] = 0; }
  int nchunk = 0;
  for (int jc = j0; jc < j1; jc += JC) {
    sf_tables<JC, BLK>(Xf, Yf, Zf, KXT, KYT, KZT, jc, j1, nq, qanc, qbase);
    __syncthreads();
    if (NA > 0) {
#pragma unroll 4
      for (int s = 0; s < JC; s += 4) {
        const float a0 = Xf[xo[0] + 2 * s] * Yf[yo[0] + 2 * s];
        const float a1 = Xf[xo[1] + 2 * s] * Yf[yo[1] + 2 * s];
        float b[NA > 0 ? NA : 1];
#pragma unroll
        for (int t = 0; t < NA; t++) b[t] = Zf[zo + 8 * ts * t * S + 2 * s];
#pragma unroll
        for (int t = 0; t < NA; t++) {
          acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b[t], acc[0][t], 0, 0, 0);
          acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b[t], acc[1][t], 0, 0, 0);
        }
      }
      // the matrix sums of 2 JC charges go to the second-level sums
      if (++nchunk % 2 == 0 || jc + JC >= j1) {
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
          for (int t = 0; t < NA; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) { mid[m][t][r] += acc[m][t][r]; acc[m][t][r] = 0.0f; }
      }
    }
    __syncthreads();
  }
  // lane l: row r0 + 4m + l/16, kz 8t + i16/2, T index 2r + i16%2
#pragma unroll
  for (int m = 0; m < 2; m++) {
    const int row = r0 + 4 * m + kq;
#pragma unroll
    for (int t = 0; t < NA; t++) {
      const int T = t0 + ts * t;
      if (row >= blk_rows[T]) continue;
      const int seg = blk_seg0[T] + row / 64;
      double *out = kpart + (((size_t)blockIdx.y * nseg + seg) * 64 + row % 64) * 64 +
                    (i16 >> 1) * 8 + (i16 & 1);
#pragma unroll
      for (int r = 0; r < 4; r++) out[2 * r] = (double)mid[m][t][r];
    }
  }
}

// WPU waves per unit of 8 rows (4 units per workgroup of 256 WPU
// threads): the waves of a unit share its A operands and take every
// WPU-th N tile, so the tables of a chunk serve 4 WPU waves and a wave
// holds NTMAX/WPU tiles' accumulators
template <int NTMAX, int JC, int WPU>
__global__ __launch_bounds__(kBlock * WPU, WPU == 1 ? 3 : 4) void k_sfac_mf(
    int nq, const double4 *__restrict__ qpos, int nrows, const int *__restrict__ row_kx,
    const int *__restrict__ row_ky, const int *__restrict__ row_kzmax,
    const int *__restrict__ blk_rows, const int *__restrict__ blk_seg0, int nseg, int KXT,
    int KYT, int KZT, int jper, double *__restrict__ kpart, const float2 *__restrict__ qanc,
    const float2 *__restrict__ qbase) {
  extern __shared__ char smem[];
  constexpr int S = SfTab<JC>::S;
  constexpr int BLK = kBlock * WPU;
  float *Xf = (float *)smem, *Yf = Xf + KXT * S, *Zf = Yf + KYT * S;
  const int wv = threadIdx.x >> 6, hf = wv % WPU;
  const int r0 = (blockIdx.x * 4 + wv / WPU) * 8;
  // N tiles of the unit (wave-uniform), 0 for a wave past the rows; this
  // wave's share: tiles hf, hf + WPU, ...
  int nt = 0;
  if (r0 < nrows) nt = min((row_kzmax[r0] + 8) / 8, NTMAX);
  nt = __builtin_amdgcn_readfirstlane(nt);
  nt = (nt - hf + WPU - 1) / WPU;
  const int j0 = blockIdx.y * jper, j1 = min(nq, j0 + jper);
#define MDX_U(NA)                                                                               \
  sf_unit<NA, JC, BLK>(Xf, Yf, Zf, KXT, KYT, KZT, j0, j1, r0, nrows, row_kx, row_ky, blk_rows,  \
                       blk_seg0, nseg, kpart, nq, qanc, qbase, hf, WPU)
  constexpr int NW = (NTMAX + WPU - 1) / WPU;   // most tiles of one wave
  switch (nt) {
    case 0: MDX_U(0); break;
    case 1: MDX_U(1); break;
    case 2: if constexpr (NW >= 2) MDX_U(2); break;
    case 3: if constexpr (NW >= 3) MDX_U(3); break;
    case 4: if constexpr (NW >= 4) MDX_U(4); break;
    case 5: if constexpr (NW >= 5) MDX_U(5); break;
    case 6: if constexpr (NW >= 6) MDX_U(6); break;
    case 7: if constexpr (NW >= 7) MDX_U(7); break;
    default: if constexpr (NW >= 8) MDX_U(8); break;
  }
#undef MDX_U
}

// Pass 1 with split-FP16 operands (split16, mdx_ksplit.h): the GEMM of k_sfac_mf
// on the FP16 instruction v_mfma_f32_16x16x32_f16 (K = 32 charges), one
// wave per unit of 8 rows with all its N tiles, U units per workgroup.
// Per chunk of JC charges the workgroup builds the cos and sin rows of X
// and Y in FP32 (rows [2k + p][charge], q folded into X) and those of Z as
// NPART FP16 parts scaled by 2^14 ([part][2 kz + r][charge]); a wave then
// forms its A operands, 8 charges of q X_p Y_q per lane for each M tile,
// scaled by asc and split, and reads its B operands from the Z parts.
// Lane l of the instruction holds A[l%16][8(l/16)..+8) and B[8(l/16)..+8)
// [l%16] (the same 8 charges for A and B) and receives D[4(l/16) + r]
// [l%16] as k_sfac_mf.  Sums: FP32 per 64 charges, FP32 second level per
// 128, FP64 per split, scaled back by 2^-14/asc.
template <int NPART>
__device__ inline void split16(float x, _Float16 (&p)[NPART]);

typedef _Float16 half2p1 __attribute__((ext_vector_type(2)));

// the three-part split of two values at once by the packed conversion
// (truncating: the remainder x - x_0 has at most 13 bits and x - x_0 - x_1
// at most 2, so every remainder is exact in FP32 and the three parts hold
// all 24 bits).  Two parts use split16 (rounded to nearest: after a
// truncated first part the second would lose two bits instead of one).
template <int NPART>
__device__ inline void split16x2(float x0, float x1, half2p1 (&p)[NPART]) {
  static_assert(NPART == 3, "the packed split is for three parts");
  float r0 = x0, r1 = x1;
#pragma unroll
  for (int k = 0; k < NPART; k++) {
    p[k] = __builtin_bit_cast(half2p1, __builtin_amdgcn_cvt_pkrtz(r0, r1));
    r0 -= (float)p[k][0];
    r1 -= (float)p[k][1];
  }
}

template <int JC>
struct SfTabH {
  static constexpr int SX = JC + 8;     // floats per X / Y row
  static constexpr int SZ = JC + 8;     // halves per Z row
};

template <int JC, int BLK, int NPART>
__device__ inline void sf_tables_h(float *Xf, float *Yf, _Float16 *Zh, int KXT, int KYT,
                                   int KZT, int jc, int j1, int nq,
                                   const float2 *__restrict__ qanc,
                                   const float2 *__restrict__ qbase, float asc) {
  constexpr int SX = SfTabH<JC>::SX, SZ = SfTabH<JC>::SZ;
  const int ntx = (KXT + 7) / 8, nty = (KYT + 7) / 8, ntz = (KZT + 7) / 8, ntask = ntx + nty + ntz;
  for (int t = threadIdx.x; t < JC * ntask; t += BLK) {
    const int jj = t % JC, task = t / JC;
    int d, m, KT;
    if (task < ntx) { d = 0; m = task; KT = KXT; }
    else if (task < ntx + nty) { d = 1; m = task - ntx; KT = KYT; }
    else { d = 2; m = task - ntx - nty; KT = KZT; }
    const int j = jc + jj;
    const int k1 = min(8 * m + 8, KT);
    float ec = 0, es = 0, bc = 0, bs = 0;
    if (j < j1) {
      const float2 a = qanc[(size_t)task * nq + j], b = qbase[(size_t)d * nq + j];
      ec = a.x; es = a.y; bc = b.x; bs = b.y;
    }
    for (int k = 8 * m; k < k1; k++) {
      if (d < 2) {
        // with three parts X carries the operand scale asc of A = q X Y (a
        // power of two); with two, the wave scales A itself
        float *dst = (d == 0 ? Xf : Yf) + 2 * k * SX + jj;
        const float sc = d == 0 && NPART == 3 ? asc : 1.0f;
        dst[0] = ec * sc;
        dst[SX] = es * sc;
      } else {
        _Float16 pc[NPART], ps[NPART];
        split16<NPART>(ec * 16384.0f, pc);
        split16<NPART>(es * 16384.0f, ps);
#pragma unroll
        for (int q = 0; q < NPART; q++) {
          Zh[(q * KZT + k) * 2 * SZ + jj] = pc[q];
          Zh[((q * KZT + k) * 2 + 1) * SZ + jj] = ps[q];
        }
      }
      const float nc = ec * bc - es * bs;
      es = ec * bs + es * bc;
      ec = nc;
    }
  }
}

typedef _Float16 half8p1 __attribute__((ext_vector_type(8)));

template <int NA, int MT, int JC, int BLK, int NPART, bool DB>
__device__ inline void sf_unit_h(char *smem, size_t set_bytes, int KXT, int KYT, int KZT, int j0,
                                 int j1, int r0, int nrows, const int *__restrict__ row_kx,
                                 const int *__restrict__ row_ky, const int *__restrict__ blk_rows,
                                 const int *__restrict__ blk_seg0, int nseg,
                                 double *__restrict__ kpart, int nq,
                                 const float2 *__restrict__ qanc,
                                 const float2 *__restrict__ qbase, float asc) {
  constexpr int SX = SfTabH<JC>::SX, SZ = SfTabH<JC>::SZ;
  const int lane = threadIdx.x & 63, i16 = lane & 15, g = lane >> 4;
  // A of M tile m: row r0 + 4m + i16/4, X_p by bit 1 of i16, Y_q by bit 0;
  // B of N tile t: kz 8t + i16/2, cos/sin by bit 0; this lane's 8 charges
  // 8g..8g+7 of each 32
  int xo[MT], yo[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int row = r0 + 4 * m + (i16 >> 2);
    int kx = 0, ky = 0;
    if (row < nrows) { kx = row_kx[row]; ky = row_ky[row]; }
    xo[m] = (2 * kx + ((i16 >> 1) & 1)) * SX + 8 * g;
    yo[m] = (2 * ky + (i16 & 1)) * SX + 8 * g;
  }
  const int zo = (2 * (i16 >> 1) + (i16 & 1)) * SZ + 8 * g;
  typedef float mf4 __attribute__((ext_vector_type(4)));
  mf4 acc[MT][NA > 0 ? NA : 1], acch[MT][NA > 0 ? NA : 1];
  float mid[MT][NA > 0 ? NA : 1][4];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int t = 0; t < NA; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) { acc[m][t][r] = 0.0f; acch[m][t][r] = 0.0f; mid[m][t][r] = 0; }
  const float unsc = 1.0f / (asc * 16384.0f);
  // table set k (DB: two sets, the next chunk's built while this one's
  // products run, one barrier per chunk)
  auto tabs = [&](int k, float *&X, float *&Y, _Float16 *&Z) {
    X = (float *)(smem + (DB ? k : 0) * set_bytes);
    Y = X + 2 * KXT * SX;
    Z = (_Float16 *)(Y + 2 * KYT * SX);
  };
  int nchunk = 0, set = 0;
  if (DB && j0 < j1) {
    float *X, *Y;
    _Float16 *Z;
    tabs(0, X, Y, Z);
    sf_tables_h<JC, BLK, NPART>(X, Y, Z, KXT, KYT, KZT, j0, j1, nq, qanc, qbase, asc);
    __syncthreads();
  }
  for (int jc = j0; jc < j1; jc += JC) {
    float *Xc, *Yc;
    _Float16 *Zc;
    tabs(set, Xc, Yc, Zc);
    const float *Xf = Xc, *Yf = Yc;
    const _Float16 *Zh = Zc;
    if (DB) {
      // the other set was last read before the previous barrier
      if (jc + JC < j1) {
        float *X, *Y;
        _Float16 *Z;
        tabs(set ^ 1, X, Y, Z);
        sf_tables_h<JC, BLK, NPART>(X, Y, Z, KXT, KYT, KZT, jc + JC, j1, nq, qanc, qbase, asc);
      }
    } else {
      sf_tables_h<JC, BLK, NPART>(Xc, Yc, Zc, KXT, KYT, KZT, jc, j1, nq, qanc, qbase, asc);
      __syncthreads();
    }
    if (NA > 0) {
      for (int s = 0; s < JC; s += 32) {
        half8p1 a[MT][NPART];
#pragma unroll
        for (int m = 0; m < MT; m++) {
          const float4 x0 = *(const float4 *)(Xf + xo[m] + s), x1 = *(const float4 *)(Xf + xo[m] + s + 4);
          const float4 y0 = *(const float4 *)(Yf + yo[m] + s), y1 = *(const float4 *)(Yf + yo[m] + s + 4);
          const float v[8] = {x0.x * y0.x, x0.y * y0.y, x0.z * y0.z, x0.w * y0.w,
                              x1.x * y1.x, x1.y * y1.y, x1.z * y1.z, x1.w * y1.w};
          if constexpr (NPART == 3) {
#pragma unroll
            for (int i = 0; i < 8; i += 2) {
              half2p1 pp[NPART];
              split16x2<NPART>(v[i], v[i + 1], pp);
#pragma unroll
              for (int q = 0; q < NPART; q++) { a[m][q][i] = pp[q][0]; a[m][q][i + 1] = pp[q][1]; }
            }
          } else {
#pragma unroll
            for (int i = 0; i < 8; i++) {
              _Float16 pp[NPART];
              split16<NPART>(v[i] * asc, pp);
#pragma unroll
              for (int q = 0; q < NPART; q++) a[m][q][i] = pp[q];
            }
          }
        }
#pragma unroll
        for (int t = 0; t < NA; t++) {
          half8p1 b[NPART];
#pragma unroll
          for (int q = 0; q < NPART; q++)
            b[q] = *(const half8p1 *)(Zh + (q * KZT + 8 * t) * 2 * SZ + zo + s);
          // products of parts (ia, ib): the leading one (0, 0) into acch,
          // the smaller ones into acc (the instruction drops low bits of
          // addends far below its accumulator, so terms of like size
          // share one); the two M tiles' chains interleaved
          constexpr int pa[6] = {NPART == 3 ? 2 : 1, 1, 0, 1, 0, 0};
          constexpr int pb[6] = {0, NPART == 3 ? 1 : 0, NPART == 3 ? 2 : 1, 0, 1, 0};
#pragma unroll
          for (int k = (NPART == 3 ? 0 : 3); k < 5; k++)
#pragma unroll
            for (int m = 0; m < MT; m++)
              acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[m][pa[k]], b[pb[k]], acc[m][t], 0, 0, 0);
#pragma unroll
          for (int m = 0; m < MT; m++)
            acch[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[m][0], b[0], acch[m][t], 0, 0, 0);
        }
      }
      // the matrix sums of 2 JC charges go to the second-level sums
      if (++nchunk % (128 / JC) == 0 || jc + JC >= j1) {
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
          for (int t = 0; t < NA; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              mid[m][t][r] += (acch[m][t][r] + acc[m][t][r]) * unsc;
              acc[m][t][r] = 0.0f;
              acch[m][t][r] = 0.0f;
            }
      }
    }
    __syncthreads();
    if (DB) set ^= 1;
  }
  // lane l: row r0 + 4m + l/16, kz 8t + i16/2, T index 2r + i16%2
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int row = r0 + 4 * m + g;
#pragma unroll
    for (int t = 0; t < NA; t++) {
      if (row >= blk_rows[t]) continue;
      const int seg = blk_seg0[t] + row / 64;
      double *out = kpart + (((size_t)blockIdx.y * nseg + seg) * 64 + row % 64) * 64 +
                    (i16 >> 1) * 8 + (i16 & 1);
#pragma unroll
      for (int r = 0; r < 4; r++) out[2 * r] = (double)mid[m][t][r];
    }
  }
}

// U units of 4 MT rows per workgroup (one wave each); NTMAX: most N tiles
template <int NTMAX, int MT, int JC, int U, int NPART, bool DB>
__global__ __launch_bounds__(64 * U, 8 / U * (MT == 1 ? 4 : 2) > 0 ? 8 / U * (MT == 1 ? 4 : 2) : 1) void k_sfac_h(
    int nq, int nrows, const int *__restrict__ row_kx, const int *__restrict__ row_ky,
    const int *__restrict__ row_kzmax, const int *__restrict__ blk_rows,
    const int *__restrict__ blk_seg0, int nseg, int KXT, int KYT, int KZT, int jper,
    double *__restrict__ kpart, const float2 *__restrict__ qanc, const float2 *__restrict__ qbase,
    float asc) {
  extern __shared__ __align__(16) char smem[];
  constexpr int SX = SfTabH<JC>::SX, SZ = SfTabH<JC>::SZ;
  constexpr int BLK = 64 * U;
  const size_t set_bytes = sizeof(float) * 2 * (KXT + KYT) * SX + sizeof(_Float16) * NPART * KZT * 2 * SZ;
  const int wv = threadIdx.x >> 6;
  const int r0 = (blockIdx.x * U + wv) * 4 * MT;
  int nt = 0;
  if (r0 < nrows) nt = min((row_kzmax[r0] + 8) / 8, NTMAX);
  nt = __builtin_amdgcn_readfirstlane(nt);
  const int j0 = blockIdx.y * jper, j1 = min(nq, j0 + jper);
#define MDX_U(NA)                                                                              \
  sf_unit_h<NA, MT, JC, BLK, NPART, DB>(smem, set_bytes, KXT, KYT, KZT, j0, j1, r0, nrows,     \
                                row_kx, row_ky, blk_rows, blk_seg0, nseg, kpart, nq, qanc,     \
                                qbase, asc)
  switch (nt) {
    case 0: MDX_U(0); break;
    case 1: MDX_U(1); break;
    case 2: if constexpr (NTMAX >= 2) MDX_U(2); break;
    case 3: if constexpr (NTMAX >= 3) MDX_U(3); break;
    case 4: if constexpr (NTMAX >= 4) MDX_U(4); break;
    case 5: if constexpr (NTMAX >= 5) MDX_U(5); break;
    case 6: if constexpr (NTMAX >= 6) MDX_U(6); break;
    case 7: if constexpr (NTMAX >= 7) MDX_U(7); break;
    default: if constexpr (NTMAX >= 8) MDX_U(8); break;
  }
#undef MDX_U
}

// sum of the charge splits: tsum[i] = sum_s kpart[s*n + i]
__global__ void k_ksum(const double *__restrict__ kpart, int nsplit, size_t n,
                       double *__restrict__ tsum) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int k = 0; k < nsplit; k++) s += kpart[(size_t)k * n + i];
  tsum[i] = s;
}

// ------------------------------------------------- energy, tensor, coefs
// one thread per (segment, lane, kz in block).  Energy and tensor follow
// sim/ewpass2m.c:194-222 written with the T sums; coefficients for pass 2:
// [0..8): 4 w e T_pq{c,s}, [8..16): kz/Lz * 4 w e (T_pqs, -T_pqc).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_kcoef(
    int nseg, const int *__restrict__ seg_blk, const int *__restrict__ seg_row0,
    const int *__restrict__ seg_rowend, const int *__restrict__ row_kx,
    const int *__restrict__ row_ky, const int *__restrict__ row_kzmax,
    const int *__restrict__ row_coef0, const double *__restrict__ tsum, double Lx, double Ly,
    double Lz, double alpha, R *__restrict__ coef, double *__restrict__ part, int measure,
    const int *__restrict__ grp_start, const int *__restrict__ grp_base,
    float *__restrict__ coefg) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int seg = t >> 9, lane = (t >> 3) & 63, k = t & 7;
  double E = 0, pk[6] = {0, 0, 0, 0, 0, 0};
  if (seg < nseg) {
    const int row = seg_row0[seg] + lane;
    const int kz = 8 * seg_blk[seg] + k;
    if (row < seg_rowend[seg] && kz <= row_kzmax[row]) {
      const int kx = row_kx[row], ky = row_ky[row];
      R *cf = coef + (size_t)(row_coef0[row] + kz) * 16;
      if (kx == 0 && ky == 0 && kz == 0) {
        for (int m = 0; m < 16; m++) cf[m] = 0;   // (the group block stays zero)
      } else {
        const double *T = tsum + ((size_t)seg * 64 + lane) * 64 + k * 8;
        const int nz = (kx > 0) + (ky > 0) + (kz > 0);
        const double w = (double)(1 << (nz - 1)) * 0.25;
        const double kxl = kx / Lx, kyl = ky / Ly, kzl = kz / Lz;
        const double kk = kxl * kxl + kyl * kyl + kzl * kzl;
        const double PIaq = -(kPi / alpha) * (kPi / alpha);
        const double e = 4 / (kk * (Lx * Ly * Lz)) * exp(PIaq * kk);
        const double we4 = 4 * w * e;
        for (int m = 0; m < 4; m++) {
          cf[2 * m] = (R)(we4 * T[2 * m]);
          cf[2 * m + 1] = (R)(we4 * T[2 * m + 1]);
          cf[8 + 2 * m] = (R)(kzl * we4 * T[2 * m + 1]);
          cf[8 + 2 * m + 1] = (R)(-kzl * we4 * T[2 * m]);
        }
        if (coefg) {
          // the same 16 values in the operand order of k_kforce_mfma: row
          // q = ky % 4 of group ky / 4, output p, cos/sin kk -> lane
          // 16(q/2) + 8(p/4) + 4(q%2) + p%4 + 32 kk of kz slot kz
          const int q = ky & 3;
          float *g = coefg + grp_base[grp_start[kx] + (ky >> 2)] + (size_t)kz * 64;
          const int ib = 16 * (q >> 1) + 4 * (q & 1);
          for (int e = 0; e < 16; e++) {
            const int pp = e >> 1, kk = e & 1;
            g[ib + 8 * (pp >> 2) + (pp & 3) + 32 * kk] = (float)cf[e];
          }
        }
        if (measure) {
          double s2 = 0;
          for (int m = 0; m < 8; m++) s2 += T[m] * T[m];
          const double ptf = 1 / kk - PIaq;
          const double E1 = w * e * 4 * s2;
          // sum_{s,t} st|Q|^2, sum t|Q|^2, sum s|Q|^2 in terms of T
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
      part[(size_t)blockIdx.x * mdx::ACC_N + mdx::ACC_EK + threadIdx.x] += s;
    }
  }
}

// dipole moment M = sum q r of the unwrapped sites (sim/ewald.c:644) and
// the second moments S_ab = sum q r_a r_b (xx yy zz yz zx xy) that the
// virial of the charged-background field needs (sim/ewald.c:909-931)
__global__ __launch_bounds__(kBlock) void k_dipole(int nq, const double4 *__restrict__ qpos,
                                                   double *__restrict__ part) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (j < nq) {
    const double4 p = qpos[j];
    m[0] = p.w * p.x; m[1] = p.w * p.y; m[2] = p.w * p.z;
    m[3] = m[0] * p.x; m[4] = m[1] * p.y; m[5] = m[2] * p.z;
    m[6] = m[1] * p.z; m[7] = m[2] * p.x; m[8] = m[0] * p.y;
  }
  __shared__ double red[kBlock / 64][9];
  const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int d = 0; d < 9; d++) {
    const double s = wave_sum(m[d]);
    if (ln == 0) red[wv][d] = s;
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++) s += red[w][threadIdx.x];
    part[(size_t)blockIdx.x * mdx::ACC_N + mdx::ACC_M0 + threadIdx.x] += s;
  }
}

// sums of the accumulators a0 + blockIdx.x over the first nblocks partial
// rows (one workgroup per accumulator, fixed order)
__global__ void k_reduce_cols(const double *part, int nblocks, int a0, double *acc) {
  __shared__ double red[kBlock];
  const int a = a0 + blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) s += part[(size_t)b * mdx::ACC_N + a];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[a] = red[0];
}

// ------------------------------------------------------------------ pass 2
// Lane <-> charge.  The kz range is taken in chunks of KZC: per chunk the
// cos/sin of c for its kz live in registers (2 KZC of them, so the chunk
// sets the register footprint and the occupancy), the rows (kx,ky) that
// reach into the chunk are visited in lexicographic order, cos/sin of a
// come from FP64 sincospi at each new kx and those of b from an FP64
// recursion along ky.  The 16 coefficients of a triple are uniform across
// the wave (scalar loads).  G, H accumulate over kz in R, forces in FP64.
// Adds -q_j grad Psi (+ dipole field) to f.
template <typename R, int KZC>
__global__ __launch_bounds__(kBlock) void k_kforce(
    int nq, const double4 *__restrict__ qpos, double iLx, double iLy, double iLz, int kxmax,
    const int4 *__restrict__ parts, const int *__restrict__ chunk_kymax,
    const int *__restrict__ lex_row0, const int2 *__restrict__ lex_meta,
    const R *__restrict__ coef, double *__restrict__ kpf) {
  constexpr int BPC = KZC / kKB;  // coefficient blocks per chunk
  const int j = blockIdx.x * kBlock + threadIdx.x;
  const int4 part = parts[blockIdx.y];   // chunk, kx begin, kx end
  const int ch = part.x;
  const double4 qp = qpos[min(j, nq - 1)];
  const double sx = frac1(qp.x * iLx), sy = frac1(qp.y * iLy), sz = frac1(qp.z * iLz);
  const R kxs = (R)iLx, kys = (R)iLy;
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
  const int *kym_c = chunk_kymax + ch * (kxmax + 1);
  double Fx = 0, Fy = 0, Fz = 0;
  for (int kx = part.y; kx < part.z; kx++) {
    const int kym = kym_c[kx];
    double ca, sa;
    cs2pi(kx * sx, ca, sa);
    const R rca = (R)ca, rsa = (R)sa;
    const R kxl = (R)kx * kxs;
    const int rbase = lex_row0[kx];
    double cb = 1, sb = 0;
    for (int ky = 0; ky <= kym; ky++) {
      const int2 mt = lex_meta[rbase + ky];
      // 4-kz blocks of this row inside the chunk (>= 1), wave-uniform
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
      const R kyl = (R)ky * kys;
      Fx += (double)(kxl * (rca * (rcb * G2 + rsb * G3) - rsa * (rcb * G0 + rsb * G1)));
      Fy += (double)(kyl * (rca * (rcb * G1 - rsb * G0) + rsa * (rcb * G3 - rsb * G2)));
      Fz += (double)(rca * (rcb * H0 + rsb * H1) + rsa * (rcb * H2 + rsb * H3));
      const double nc = cb * cb1 - sb * sb1;
      sb = cb * sb1 + sb * cb1;
      cb = nc;
    }
  }
  if (j >= nq) return;
  double *o = kpf + ((size_t)blockIdx.y * nq + j) * 3;
  o[0] = Fx;
  o[1] = Fy;
  o[2] = Fz;
}

// Pass 2 on the matrix cores (MDX_MIXED).  For a group of 4 rows (kx,
// ky0..ky0+3) of the chunk and 32 charges, the 8 sums G0..3, H0..3 of each
// row and charge are one product D[32 x 32] = E[32 x 2 KZC] Z[2 KZC x 32]
// with E = the rows' coefficients and Z = the charges' cos/sin of c along
// kz: a chain of v_mfma_f32_32x32x2_f32, one per kz (K = cos, sin).
// Layout of that instruction on gfx950 (checked on the device): lane l
// holds A[l%32][l/32], B[l/32][l%32] and, in register r, D[4(l/32) + r%4 +
// 8(r/4)][l%32].  The M index i of D is mapped to row 2(i/16) + (i/4)%2
// and output 4((i/8)%2) + i%4, so that lane l (charge l%32, half h = l/32)
// receives all 8 outputs of rows h and 2+h of the group in registers 0-7
// and 8-15, and forms their x, y and z forces itself; it supplies the cos
// (h = 0) or sin (h = 1) of its charge along kz and one coefficient of one
// row.  Rows past the chunk's last ky of a kx and kz past a row's last
// block contribute zeros.  Row phases: FP64 anchors per kx, FP64 steps of
// two rows along ky.  The products and sums are FP32 as in k_kforce (on
// gfx950 the results are the same bits).
typedef float mfma_f32x16 __attribute__((ext_vector_type(16)));

template <int KZC, int NTC>
__global__ __launch_bounds__(kBlock, NTC == 1 ? 4 : 3) void k_kforce_mfma(
    int nq, const double4 *__restrict__ qpos, double iLx, double iLy, double iLz, int kxmax,
    const int4 *__restrict__ parts, const int *__restrict__ chunk_kymax,
    const int *__restrict__ grp_start, const int *__restrict__ grp_base,
    const int *__restrict__ grp_nb, const float *__restrict__ coefg, double *__restrict__ kpf) {
  // NTC tiles of 32 charges per wave share each coefficient operand: the
  // products of tile u go to acc[u]
  constexpr int BPC = KZC / kKB;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = lane & 31, h = lane >> 5;
  const int jw = (blockIdx.x * (kBlock / 64) + wv) * 32 * NTC + c;
  const int4 part = parts[blockIdx.y];
  const int ch = part.x;
  const float kxs = (float)iLx, kys = (float)iLy;
  double sx[NTC], cb1[NTC], sb1[NTC], cb2[NTC], sb2[NTC], cb4[NTC], sb4[NTC];
  float zt[NTC][KZC];
#pragma unroll
  for (int u = 0; u < NTC; u++) {
    const double4 qp = qpos[min(jw + 32 * u, nq - 1)];
    sx[u] = frac1(qp.x * iLx);
    const double sy = frac1(qp.y * iLy), sz = frac1(qp.z * iLz);
    // e^{2 pi i sy} to the powers 1, 2, 4
    cs2pi(sy, cb1[u], sb1[u]);
    cb2[u] = cb1[u] * cb1[u] - sb1[u] * sb1[u];
    sb2[u] = 2 * cb1[u] * sb1[u];
    cb4[u] = cb2[u] * cb2[u] - sb2[u] * sb2[u];
    sb4[u] = 2 * cb2[u] * sb2[u];
    // this lane's half of the kz table: cos (h = 0) or sin (h = 1)
    double cz1, sz1;
    cs2pi(sz, cz1, sz1);
#pragma unroll
    for (int m = 0; m < KZC / 8; m++) {
      double cc, ss;
      cs2pi((double)(ch * KZC + 8 * m) * sz, cc, ss);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        zt[u][8 * m + k] = (float)(h ? ss : cc);
        const double nc = cc * cz1 - ss * sz1;
        ss = cc * sz1 + ss * cz1;
        cc = nc;
      }
    }
  }
  const int *kym_c = chunk_kymax + ch * (kxmax + 1);
  const float *cfl = coefg + lane + (size_t)ch * KZC * 64;
  double Fx[NTC], Fy[NTC], Fz[NTC];
#pragma unroll
  for (int u = 0; u < NTC; u++) { Fx[u] = 0; Fy[u] = 0; Fz[u] = 0; }
  for (int kx = part.y; kx < part.z; kx++) {
    const int kym = kym_c[kx];
    float rca[NTC], rsa[NTC];
    double cb[NTC], sb[NTC];
#pragma unroll
    for (int u = 0; u < NTC; u++) {
      double ca, sa;
      cs2pi(kx * sx[u], ca, sa);
      rca[u] = (float)ca; rsa[u] = (float)sa;
      // phase of row ky0 + h
      cb[u] = h ? cb1[u] : 1.0; sb[u] = h ? sb1[u] : 0.0;
    }
    const float kxl = (float)kx * kxs;
    const int g0 = grp_start[kx];
    for (int ky0 = 0; ky0 <= kym; ky0 += 4) {
      const int gi = g0 + (ky0 >> 2);
      const int nb = min(grp_nb[gi] - ch * BPC, BPC);   // wave-uniform, >= 1
      const float *ga = cfl + grp_base[gi];
      mfma_f32x16 acc[NTC];
#pragma unroll
      for (int u = 0; u < NTC; u++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[u][r] = 0.0f;
#pragma unroll
      for (int b = 0; b < BPC; b++) {
        if (b < nb) {
          float av[kKB];
#pragma unroll
          for (int k = 0; k < kKB; k++) av[k] = ga[(b * kKB + k) * 64];
#pragma unroll
          for (int k = 0; k < kKB; k++)
#pragma unroll
            for (int u = 0; u < NTC; u++)
              acc[u] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[k], zt[u][b * kKB + k], acc[u], 0, 0, 0);
        }
      }
      // rows ky0 + h (registers 0-7) and ky0 + 2 + h (8-15); rows past kym
      // hold zeros
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
          const float rcb = (float)cq, rsb = (float)sq;
          const float kyl = (float)(ky0 + 2 * q + h) * kys;
          Fx[u] += (double)(kxl * (rca[u] * (rcb * G2 + rsb * G3) - rsa[u] * (rcb * G0 + rsb * G1)));
          Fy[u] += (double)(kyl * (rca[u] * (rcb * G1 - rsb * G0) + rsa[u] * (rcb * G3 - rsb * G2)));
          Fz[u] += (double)(rca[u] * (rcb * H0 + rsb * H1) + rsa[u] * (rcb * H2 + rsb * H3));
        }
        const double nc = cb[u] * cb4[u] - sb[u] * sb4[u];
        sb[u] = cb[u] * sb4[u] + sb[u] * cb4[u];
        cb[u] = nc;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NTC; u++) {
    // the two halves hold the even and odd rows of the same charges
    const double fx = Fx[u] + __shfl_xor(Fx[u], 32, 64);
    const double fy = Fy[u] + __shfl_xor(Fy[u], 32, 64);
    const double fz = Fz[u] + __shfl_xor(Fz[u], 32, 64);
    const int j = jw + 32 * u;
    if (h || j >= nq) continue;
    double *o = kpf + ((size_t)blockIdx.y * nq + j) * 3;
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
  }
}

// Pass 2 with split-FP16 operands (kforce_h_body, mdx_ksplit.h): grid
// (blocks of 32 NTC kBlock / 64 charges, parts)
template <int KZC, int NTC, int NPART>
__global__ __launch_bounds__(kBlock, NTC == 1 ? 3 : 2) void k_kforce_h(
    int nq, const double4 *__restrict__ qpos, double iLx, double iLy, double iLz, int kxmax,
    const int4 *__restrict__ parts, const int *__restrict__ chunk_kymax,
    const int *__restrict__ grp_start, const int *__restrict__ grp_hb,
    const int *__restrict__ grp_nb8, const half8 *__restrict__ coefh,
    const float *__restrict__ gscale, double *__restrict__ kpf) {
  kforce_h_body<KZC, NTC, NPART>(nq, qpos, iLx, iLy, iLz, kxmax, parts[blockIdx.y], blockIdx.x,
                                 chunk_kymax, grp_start, grp_hb, grp_nb8, coefh, gscale,
                                 kpf + (size_t)blockIdx.y * nq * 3);
}

// the cell-dipole field on a charge at r: ek_d = -cm_d M_d + cb_d r_d
struct DipField {
  double cm[3], cb[3];
};

// sums the parts of pass 2 in a fixed order and adds -q grad Psi and the
// cell-dipole field to the force rows of the charges
__global__ void k_kforce_sum(int nq, int nparts, const double *__restrict__ kpf,
                             const int *__restrict__ qsite, const double4 *__restrict__ qpos,
                             const double *__restrict__ acc, DipField df,
                             double *__restrict__ f) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nq) return;
  double F[3] = {0, 0, 0};
  for (int p = 0; p < nparts; p++) {
    const double *o = kpf + ((size_t)p * nq + j) * 3;
    F[0] += o[0];
    F[1] += o[1];
    F[2] += o[2];
  }
  const int site = qsite[j];   // row of f (Drudes after the centres)
  const double4 p = qpos[j];
  const double q = p.w, R[3] = {p.x, p.y, p.z};
  const double *M = acc + mdx::ACC_M0;
  for (int d = 0; d < 3; d++)
    f[3 * site + d] += -q * F[d] + q * (-df.cm[d] * M[d] + df.cb[d] * R[d]);
}

template <typename R>
int run_kspace(mdx::Ctx &c, int measure, hipStream_t st) {
  using namespace mdx;
  const KLayout &K = c.K;
  const int KXT = K.kxmax + 1, KYT = K.row_ky.empty() ? 1 : *std::max_element(K.row_ky.begin(), K.row_ky.end()) + 1;
  const int KZT = K.kzpad;
  const size_t shmem = sizeof(C2<R>) * kJC * (KXT + KYT + KZT);
  const int jper = (c.nq + c.nsplit - 1) / c.nsplit;
  dim3 g1((K.nseg + 3) / 4, c.nsplit);
  const double4 *qpos = (const double4 *)c.qpos.p;
  // anchors of the matrix-core pass 1 (mixed mode, row units)
  const bool anc = sizeof(R) == 4 && (c.p1nt == 1 || c.p1nt == 2 || (c.p1nt >= 4 && c.p1nt <= 7));
  const int3 nb = make_int3((KXT + 7) / 8, (KYT + 7) / 8, (KZT + 7) / 8);
  const int3 boff = make_int3(0, nb.x, nb.x + nb.y);
  if (anc && (size_t)(nb.x + nb.y + nb.z + 3) * c.nq > c.qanc.n)
    return fail(MDX_E_ARG, "k-space: anchor buffer not set up");
  k_qgather<<<nblk(c.nq), kBlock, 0, st>>>(c.nq, c.ns, c.q_site.p, c.r.p, c.charge.p,
                                          c.polar ? c.chargepol.p : nullptr, 1 / c.L[0],
                                          1 / c.L[1], 1 / c.L[2], (double4 *)c.qpos.p,
                                          (double4 *)c.qtrig.p, anc ? c.qanc.p : nullptr,
                                          anc ? c.qanc.p + (size_t)(nb.x + nb.y + nb.z) * c.nq
                                              : nullptr,
                                          nb, boff);
  if (c.timing) c.tm.start(T_KPASS1, st);
  bool pass1_mfma = false;
  int nsplit1 = c.nsplit;              // charge splits written to kpart
  if constexpr (sizeof(R) == 4) pass1_mfma = c.p1nt != 0;
  if (pass1_mfma && (c.p1nt == 1 || c.p1nt == 2)) {
    // row units of 8 rows, 4 units per workgroup (one or two waves each);
    // splits for about 1024 workgroups
    const int wgx = ((K.nrows + 7) / 8 + 3) / 4;
    nsplit1 = std::max(1, std::min(c.nsplit, 1024 / wgx));
    const int jp = (c.nq + nsplit1 - 1) / nsplit1;
    const dim3 gm(wgx, nsplit1);
    constexpr int JC = 64;
    const size_t shm = sizeof(float) * SfTab<JC>::S * (KXT + KYT + KZT);
#define MDX_SF(NT, WPU)                                                                       \
  k_sfac_mf<NT, JC, WPU><<<gm, kBlock * WPU, shm, st>>>(                                      \
      c.nq, qpos, K.nrows, c.d_row_kx.p, c.d_row_ky.p, c.d_row_kzmax.p, c.d_blk_rows.p,        \
      c.d_blk_seg0.p, K.nseg, KXT, KYT, KZT, jp, c.kpart.p, c.qanc.p,                          \
      c.qanc.p + (size_t)(nb.x + nb.y + nb.z) * c.nq)
    if (K.nblk > 8)
      return fail(MDX_E_ARG, "k-space pass 1 on row units: at most 8 kz blocks (kz < 64)");
    if (c.p1nt == 2) {
      if (K.nblk <= 2) MDX_SF(2, 2);
      else if (K.nblk <= 4) MDX_SF(4, 2);
      else if (K.nblk <= 6) MDX_SF(6, 2);
      else MDX_SF(8, 2);
    } else {
      if (K.nblk <= 2) MDX_SF(2, 1);
      else if (K.nblk <= 4) MDX_SF(4, 1);
      else if (K.nblk <= 6) MDX_SF(6, 1);
      else MDX_SF(8, 1);
    }
#undef MDX_SF
  } else if (pass1_mfma && c.p1nt >= 4 && c.p1nt <= 7) {
    // split-FP16 row units, one wave each: (M tiles per wave, waves per
    // workgroup, charges per chunk, table sets) = (2, 8, 64, 1),
    // (1, 16, 64, 1), (1, 16, 32, 2), (2, 8, 32, 2) for p1nt 4..7
    if (K.nblk > 8)
      return fail(MDX_E_ARG, "k-space pass 1 on row units: at most 8 kz blocks (kz < 64)");
    const int cfg = c.p1nt - 4;
    const int MTv = cfg == 0 || cfg == 3 ? 2 : 1, Uv = cfg == 1 || cfg == 2 ? 16 : 8;
    const int JCv = cfg >= 2 ? 32 : 64, nset = cfg >= 2 ? 2 : 1;
    const int wgx = ((K.nrows + 4 * MTv - 1) / (4 * MTv) + Uv - 1) / Uv;
    nsplit1 = std::max(1, std::min(c.nsplit, 1024 / wgx));
    const int jp = (c.nq + nsplit1 - 1) / nsplit1;
    const int npart = c.kprod == 3 ? 2 : 3;
    const size_t shm = nset * (sizeof(float) * 2 * (KXT + KYT) * (JCv + 8) +
                               sizeof(_Float16) * npart * KZT * 2 * (JCv + 8));
    // A = q X Y scaled into [2^13, 2^14) at the largest |q|
    const float asc = std::ldexp(1.0f, 13 - std::ilogb(std::max(c.qmax, 1e-30)));
    const dim3 gm(wgx, nsplit1);
#define MDX_SH(NT, MT, JC, U, NP)                                                            \
  k_sfac_h<NT, MT, JC, U, NP, JC == 32><<<gm, 64 * U, shm, st>>>(                            \
      c.nq, K.nrows, c.d_row_kx.p, c.d_row_ky.p, c.d_row_kzmax.p, c.d_blk_rows.p,           \
      c.d_blk_seg0.p, K.nseg, KXT, KYT, KZT, jp, c.kpart.p, c.qanc.p,                       \
      c.qanc.p + (size_t)(nb.x + nb.y + nb.z) * c.nq, asc)
#define MDX_SHN(MT, JC, U, NP)                       \
  if (K.nblk <= 2) MDX_SH(2, MT, JC, U, NP);         \
  else if (K.nblk <= 4) MDX_SH(4, MT, JC, U, NP);    \
  else if (K.nblk <= 6) MDX_SH(6, MT, JC, U, NP);    \
  else MDX_SH(8, MT, JC, U, NP);
#define MDX_SHC(NP)                                  \
  switch (cfg) {                                     \
    case 0: MDX_SHN(2, 64, 8, NP); break;            \
    case 1: MDX_SHN(1, 64, 16, NP); break;           \
    case 2: MDX_SHN(1, 32, 16, NP); break;           \
    default: MDX_SHN(2, 32, 8, NP); break;           \
  }
    if (npart == 3) { MDX_SHC(3); } else { MDX_SHC(2); }
#undef MDX_SHC
#undef MDX_SHN
#undef MDX_SH
  } else if (pass1_mfma && c.p1nt == 8)
    k_sfac_mfma<8><<<dim3((K.nseg + 1) / 2, c.nsplit), kBlock, shmem, st>>>(
        c.nq, qpos, K.nseg, c.d_seg_blk.p, c.d_seg_row0.p, c.d_seg_rowend.p, c.d_row_kx.p,
        c.d_row_ky.p, KXT, KYT, KZT, jper, c.kpart.p, (const double4 *)c.qtrig.p);
  else if (pass1_mfma)
    k_sfac_mfma<16><<<g1, kBlock, shmem, st>>>(
        c.nq, qpos, K.nseg, c.d_seg_blk.p, c.d_seg_row0.p, c.d_seg_rowend.p, c.d_row_kx.p,
        c.d_row_ky.p, KXT, KYT, KZT, jper, c.kpart.p, (const double4 *)c.qtrig.p);
  else
    k_sfac<R><<<g1, kBlock, shmem, st>>>(
        c.nq, qpos, 1 / c.L[0], 1 / c.L[1], 1 / c.L[2], K.nseg,
        c.d_seg_blk.p, c.d_seg_row0.p, c.d_seg_rowend.p, c.d_row_kx.p, c.d_row_ky.p, KXT, KYT,
        KZT, jper, c.kpart.p, (const double4 *)c.qtrig.p);
  MDX_HIP(hipGetLastError());
  if (c.timing) { c.tm.stop(T_KPASS1, st); c.tm.start(T_KCOEF, st); }
  const size_t nt = (size_t)K.nseg * 64 * 64;
  k_ksum<<<nblk(nt), kBlock, 0, st>>>(c.kpart.p, nsplit1, nt, c.tsum.p);
  R *coef = nullptr;
  if constexpr (sizeof(R) == 4) coef = c.coef32.p; else coef = (R *)c.coef64.p;
  float *coefg = nullptr;
  if constexpr (sizeof(R) == 4) coefg = c.kmfma ? c.coefg.p : nullptr;
  k_kcoef<R><<<nblk((long)K.nseg * 512), kBlock, 0, st>>>(
      K.nseg, c.d_seg_blk.p, c.d_seg_row0.p, c.d_seg_rowend.p, c.d_row_kx.p, c.d_row_ky.p,
      c.d_row_kzmax.p, c.d_row_coef0.p, c.tsum.p, c.L[0], c.L[1], c.L[2], c.alpha, coef, c.part.p,
      measure, c.d_grp_start.p, c.d_grp_base.p, coefg);
  MDX_HIP(hipGetLastError());
  if constexpr (sizeof(R) == 4) {
    if (c.kmfma && c.kprod) {
      // the group coefficients as split-FP16 operands of pass 2
      const int ngrp = (int)K.grp_base.size();
      if (c.kprod == 6)
        k_coefh<3><<<ngrp, kBlock, 0, st>>>(c.d_grp_base.p, c.d_grp_nb.p, c.d_grp_hb.p,
                                          c.d_grp_nb8.p, c.coefg.p, (half8 *)c.coefh.p,
                                          c.gscale.p);
      else
        k_coefh<2><<<ngrp, kBlock, 0, st>>>(c.d_grp_base.p, c.d_grp_nb.p, c.d_grp_hb.p,
                                          c.d_grp_nb8.p, c.coefg.p, (half8 *)c.coefh.p,
                                          c.gscale.p);
      MDX_HIP(hipGetLastError());
    }
  }
  k_dipole<<<nblk(c.nq), kBlock, 0, st>>>(c.nq, qpos, c.part.p);
  // the dipole and second-moment sums of k_dipole's workgroups
  k_reduce_cols<<<9, kBlock, 0, st>>>(c.part.p, nblk(c.nq), ACC_M0, c.acc.p);
  if (c.timing) { c.tm.stop(T_KCOEF, st); c.tm.start(T_KPASS2, st); }
  if constexpr (sizeof(R) == 4) {
    if (c.kmfma && c.kprod) {
      // split-FP16 operands (k_coefh made them beside the coefficients)
#define MDX_KH(KZV, NTV, NPV)                                                                \
  k_kforce_h<KZV, NTV, NPV><<<dim3(nblk(c.nq, 128 * NTV), c.nkpart), kBlock, 0, st>>>(      \
      c.nq, qpos, 1 / c.L[0], 1 / c.L[1], 1 / c.L[2], K.kxmax, (const int4 *)c.d_kparts.p,  \
      c.d_chunk_kymax.p, c.d_grp_start.p, c.d_grp_hb.p, c.d_grp_nb8.p,                        \
      (const half8 *)c.coefh.p, c.gscale.p, c.kpf.p)
#define MDX_KHZ(NTV, NPV)                   \
  switch (c.kzc) {                          \
    case 8: MDX_KH(8, NTV, NPV); break;     \
    case 16: MDX_KH(16, NTV, NPV); break;   \
    case 24: MDX_KH(24, NTV, NPV); break;   \
    case 32: MDX_KH(32, NTV, NPV); break;   \
    case 48: MDX_KH(48, NTV, NPV); break;   \
    case 64: MDX_KH(64, NTV, NPV); break;   \
    default: return fail(MDX_E_ARG, "k-space: kz chunk must be 8, 16, 24, 32, 48 or 64"); \
  }
      if (c.kprod == 6) {
        if (c.kmfma == 2) { MDX_KHZ(2, 3); } else { MDX_KHZ(1, 3); }
      } else {
        if (c.kmfma == 2) { MDX_KHZ(2, 2); } else { MDX_KHZ(1, 2); }
      }
#undef MDX_KHZ
#undef MDX_KH
      MDX_HIP(hipGetLastError());
      return 0;
    }
    if (c.kmfma) {
#define MDX_KM(KZV)                                                                         \
  k_kforce_mfma<KZV, KNTC><<<dim3(nblk(c.nq, 128 * KNTC), c.nkpart), kBlock, 0, st>>>(      \
      c.nq, qpos, 1 / c.L[0], 1 / c.L[1], 1 / c.L[2], K.kxmax, (const int4 *)c.d_kparts.p,  \
      c.d_chunk_kymax.p, c.d_grp_start.p, c.d_grp_base.p, c.d_grp_nb.p, c.coefg.p, c.kpf.p)
      if (c.kmfma == 2) {
        constexpr int KNTC = 2;
        switch (c.kzc) {
          case 8: MDX_KM(8); break;
          case 16: MDX_KM(16); break;
          case 24: MDX_KM(24); break;
          case 32: MDX_KM(32); break;
          case 48: MDX_KM(48); break;
          case 64: MDX_KM(64); break;
          default: return fail(MDX_E_ARG, "k-space: kz chunk must be 8, 16, 24, 32, 48 or 64");
        }
      } else {
        constexpr int KNTC = 1;
        switch (c.kzc) {
          case 8: MDX_KM(8); break;
          case 16: MDX_KM(16); break;
          case 24: MDX_KM(24); break;
          case 32: MDX_KM(32); break;
          case 48: MDX_KM(48); break;
          case 64: MDX_KM(64); break;
          default: return fail(MDX_E_ARG, "k-space: kz chunk must be 8, 16, 24, 32, 48 or 64");
        }
      }
#undef MDX_KM
      MDX_HIP(hipGetLastError());
      return 0;
    }
  }
#define MDX_KF(KZV)                                                                         \
  k_kforce<R, KZV><<<dim3(nblk(c.nq), c.nkpart), kBlock, 0, st>>>(                    \
      c.nq, qpos, 1 / c.L[0], 1 / c.L[1], 1 / c.L[2], K.kxmax, (const int4 *)c.d_kparts.p,  \
      c.d_chunk_kymax.p, c.d_lex_row0.p, (const int2 *)c.d_lex_meta.p, coef, c.kpf.p)
  switch (c.kzc) {
    case 8: MDX_KF(8); break;
    case 16: MDX_KF(16); break;
    case 24: MDX_KF(24); break;
    case 32: MDX_KF(32); break;
    case 48: MDX_KF(48); break;
    case 64: MDX_KF(64); break;
    default: return fail(MDX_E_ARG, "k-space: kz chunk must be 8, 16, 24, 32, 48 or 64");
  }
#undef MDX_KF
  MDX_HIP(hipGetLastError());
  return 0;
}

// the parts of pass 2 and the dipole field into the force rows
int add_kspace(mdx::Ctx &c, hipStream_t st) {
  using namespace mdx;
  const double V = c.L[0] * c.L[1] * c.L[2];
  DipField df;
  double ce[3];
  ewald_dipole_coefs(c.el_corr, c.el_bg, c.epsinf, c.qnet, V, df.cm, ce, df.cb);
  k_kforce_sum<<<nblk(c.nq), kBlock, 0, st>>>(c.nq, c.nkpart, c.kpf.p, c.q_site.p,
                                             (const double4 *)c.qpos.p, c.acc.p, df, c.f.p);
  MDX_HIP(hipGetLastError());
  if (c.timing) c.tm.stop(T_KPASS2, st);
  return 0;
}

}  // namespace

namespace mdx {

// k-vector set of Ewald(0) for box L and kappa (sim/ewald.c:501-534): the
// rows (kx,ky), kx in [0,(int)(L0 kappa)], ky in [0,kymax(kx)], with kz
// from !(kx+ky) to kzmax(kx,ky).  Rows are sorted by kz limit (pass 1
// segments) and also kept in lexicographic order (pass 2).
void klayout_build(const double L[3], double kappa, KLayout &K) {
  K = KLayout();
  const double k2 = kappa * kappa;
  K.kxmax = (int)(L[0] * kappa);
  std::vector<int> lx, ly, lz, lrow0(K.kxmax + 1);
  for (int kx = 0; kx <= K.kxmax; kx++) {
    const double rx = k2 - (kx / L[0]) * (kx / L[0]);
    const int kym = (int)(L[1] * std::sqrt(rx));
    lrow0[kx] = (int)lx.size();
    for (int ky = 0; ky <= kym; ky++) {
      const double ry = rx - (ky / L[1]) * (ky / L[1]);
      lx.push_back(kx);
      ly.push_back(ky);
      lz.push_back((int)(L[2] * std::sqrt(ry)));
    }
  }
  K.nrows = (int)lx.size();
  int kzmax_all = *std::max_element(lz.begin(), lz.end());
  K.nblk = kzmax_all / 8 + 1;
  K.kzpad = 8 * K.nblk;
  // coefficient offsets in lexicographic order, rows padded to kKB kz
  std::vector<int> lcoef(K.nrows), lnb(K.nrows);
  int off = 0;
  K.ntrip = 0;
  for (int i = 0; i < K.nrows; i++) {
    lcoef[i] = off;
    lnb[i] = (lz[i] + kKB) / kKB;
    off += lnb[i] * kKB;
    K.ntrip += lz[i] + 1 - (i == 0);
  }
  K.ncoef = off;
  // rows sorted by kz limit (descending), stable
  std::vector<int> ord(K.nrows);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return lz[a] > lz[b]; });
  for (int i : ord) {
    K.row_kx.push_back(lx[i]);
    K.row_ky.push_back(ly[i]);
    K.row_kzmax.push_back(lz[i]);
    K.row_coef0.push_back(lcoef[i]);
  }
  for (int b = 0; b < K.nblk; b++) {
    int n = 0;
    while (n < K.nrows && K.row_kzmax[n] >= 8 * b) n++;
    for (int r0 = 0; r0 < n; r0 += 64) {
      K.seg_blk.push_back(b);
      K.seg_row0.push_back(r0);
      K.seg_rowend.push_back(std::min(r0 + 64, n));
    }
  }
  K.nseg = (int)K.seg_blk.size();
  K.lex_row0 = lrow0;
  K.lex_meta.resize(2 * K.nrows);
  K.lex_maxblk = 0;
  for (int i = 0; i < K.nrows; i++) {
    K.lex_meta[2 * i] = lnb[i];
    K.lex_meta[2 * i + 1] = lcoef[i];
    K.lex_maxblk = std::max(K.lex_maxblk, lnb[i]);
  }
  // row groups: for each kx the rows ky = 4g..4g+3; a group's block holds
  // 4 lnb(first row) kz slots of 64 floats (32 M entries x cos/sin)
  K.grp_start.assign(K.kxmax + 2, 0);
  K.grp_floats = 0;
  for (int kx = 0; kx <= K.kxmax; kx++) {
    K.grp_start[kx] = (int)K.grp_base.size();
    const int r0 = K.lex_row0[kx], r1 = kx < K.kxmax ? K.lex_row0[kx + 1] : K.nrows;
    for (int r = r0; r < r1; r += 4) {
      K.grp_base.push_back((int)K.grp_floats);
      K.grp_nb.push_back(lnb[r]);
      K.grp_floats += (long)lnb[r] * kKB * 64;
      K.grp_hb.push_back((int)K.grp_h8);
      K.grp_nb8.push_back((lnb[r] * kKB + 7) / 8);
      K.grp_h8 += (lnb[r] * kKB + 7) / 8;
    }
  }
  K.grp_start[K.kxmax + 1] = (int)K.grp_base.size();
}

int kspace_setup(Ctx &c) {
  KLayout &K = c.K;
  klayout_build(c.L, c.kappa, K);
  // the coefficient kernel writes 2 partial rows per segment, the dipole
  // kernel one per 256 charges
  const int need = std::max(2 * K.nseg, (c.nq + kBlock - 1) / kBlock);
  if (need > c.part_blocks) {
    c.part_blocks = need;
    int rc0;
    if ((rc0 = c.part.alloc((size_t)c.part_blocks * ACC_N))) return rc0;
    c.gen++;
  }
  // charge splits: about 8 workgroups per CU, at least 256 charges each
  const int wg = (K.nseg + 3) / 4;
  c.nsplit = std::max(1, std::min(2048 / std::max(wg, 1), c.nq / 256));
  // device copies; the lexicographic tables follow the sorted ones
  auto cat = [](std::vector<int> a, const std::vector<int> &b) {
    a.insert(a.end(), b.begin(), b.end());
    return a;
  };
  std::vector<int> kx_all = K.row_kx, ky_all = K.row_ky;
  std::vector<int> kz_all = K.row_kzmax, c0_all = K.row_coef0;
  (void)cat;
  int rc;
  if ((rc = c.d_row_kx.alloc(kx_all.size())) || (rc = c.d_row_ky.alloc(ky_all.size())) ||
      (rc = c.d_row_kzmax.alloc(kz_all.size())) || (rc = c.d_row_coef0.alloc(c0_all.size())) ||
      (rc = c.d_seg_blk.alloc(K.nseg)) || (rc = c.d_seg_row0.alloc(K.nseg)) ||
      (rc = c.d_seg_rowend.alloc(K.nseg)) ||
      (rc = c.kpart.alloc((size_t)c.nsplit * K.nseg * 64 * 64)) ||
      (rc = c.tsum.alloc((size_t)K.nseg * 64 * 64)) || (rc = c.d_blk_rows.alloc(K.nblk)) ||
      (rc = c.d_blk_seg0.alloc(K.nblk)))
    return rc;
  {
    // per 8-kz block: rows reaching it and its first segment
    std::vector<int> br(K.nblk, 0), bs(K.nblk, -1);
    for (int sg = 0; sg < K.nseg; sg++) {
      const int b = K.seg_blk[sg];
      if (bs[b] < 0) bs[b] = sg;
      br[b] = std::max(br[b], K.seg_rowend[sg]);
    }
    MDX_HIP(hipMemcpy(c.d_blk_rows.p, br.data(), sizeof(int) * K.nblk, hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.d_blk_seg0.p, bs.data(), sizeof(int) * K.nblk, hipMemcpyHostToDevice));
  }
  // coefficient rows: the padding slots stay zero (k_kcoef writes kz<=kzmax)
  if (c.mode == MDX_MIXED) {
    // anchors and bases of the row-unit pass 1 (k_qgather), made here so
    // that no call allocates (a captured step cannot)
    const int KYT = K.row_ky.empty() ? 1 : *std::max_element(K.row_ky.begin(), K.row_ky.end()) + 1;
    const size_t nanc = (size_t)((K.kxmax + 8) / 8 + (KYT + 7) / 8 + (K.kzpad + 7) / 8 + 3) * c.nq;
    if ((rc = c.qanc.alloc(nanc))) return rc;
    if ((rc = c.coef32.alloc((size_t)K.ncoef * 16))) return rc;
    MDX_HIP(hipMemset(c.coef32.p, 0, sizeof(float) * K.ncoef * 16));
    // row-group blocks of the matrix-core pass 2 (padding stays zero)
    if ((rc = c.coefg.alloc((size_t)K.grp_floats)) || (rc = c.d_grp_start.alloc(K.grp_start.size())) ||
        (rc = c.d_grp_base.alloc(K.grp_base.size())) || (rc = c.d_grp_nb.alloc(K.grp_nb.size())))
      return rc;
    MDX_HIP(hipMemset(c.coefg.p, 0, sizeof(float) * K.grp_floats));
    // split-FP16 operands of the same groups (up to 3 parts); k_coefh
    // writes every slot of every call
    const size_t ngrp = K.grp_base.size();
    if ((rc = c.coefh.alloc((size_t)K.grp_h8 * 3 * 512)) || (rc = c.gscale.alloc(ngrp)) ||
        (rc = c.d_grp_hb.alloc(ngrp)) || (rc = c.d_grp_nb8.alloc(ngrp)))
      return rc;
    MDX_HIP(hipMemcpy(c.d_grp_hb.p, K.grp_hb.data(), sizeof(int) * ngrp, hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.d_grp_nb8.p, K.grp_nb8.data(), sizeof(int) * ngrp, hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.d_grp_start.p, K.grp_start.data(), sizeof(int) * K.grp_start.size(), hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.d_grp_base.p, K.grp_base.data(), sizeof(int) * K.grp_base.size(), hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.d_grp_nb.p, K.grp_nb.data(), sizeof(int) * K.grp_nb.size(), hipMemcpyHostToDevice));
  } else {
    if ((rc = c.coef64.alloc((size_t)K.ncoef * 16))) return rc;
    MDX_HIP(hipMemset(c.coef64.p, 0, sizeof(double) * K.ncoef * 16));
  }
  auto up = [&](int *dst, const std::vector<int> &v) {
    return hipMemcpy(dst, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up(c.d_row_kx.p, kx_all) || !up(c.d_row_ky.p, ky_all) || !up(c.d_row_kzmax.p, kz_all) ||
      !up(c.d_row_coef0.p, c0_all) || !up(c.d_seg_blk.p, K.seg_blk) ||
      !up(c.d_seg_row0.p, K.seg_row0) || !up(c.d_seg_rowend.p, K.seg_rowend))
    return fail(MDX_E_GPU, "kspace_setup: upload failed");
  if ((rc = c.d_lex_row0.alloc(K.lex_row0.size())) || (rc = c.d_lex_meta.alloc(K.lex_meta.size())))
    return rc;
  if (!up(c.d_lex_row0.p, K.lex_row0) || !up(c.d_lex_meta.p, K.lex_meta))
    return fail(MDX_E_GPU, "kspace_setup: upload failed");
  return kspace_chunks(c, c.kzc, c.kparts);
}

// Work split of pass 2 for kz chunk kzc into about `parts` parts: per
// chunk the largest ky of each kx whose row reaches into the chunk, and
// parts (chunk, kx begin, kx end) of about equal numbers of coefficient
// blocks (a row costs its blocks plus about two blocks of phase and
// force arithmetic).
int kchunks_build(const KLayout &K, int kzc, int parts, std::vector<int> &kym,
                  std::vector<int> &ptrim, int &nchunk_out) {
  if (kzc < 8 || kzc > 64 || kzc % 8) return fail(MDX_E_ARG, "k-space: kz chunk in {8,16,...,64}");
  if (parts < 1 || parts > 256) return fail(MDX_E_ARG, "k-space: parts in 1..256");
  const int bpc = kzc / kKB, nkx = K.kxmax + 1;
  const int nchunk = std::max(1, (K.lex_maxblk + bpc - 1) / bpc);
  kym.assign((size_t)nchunk * nkx, -1);
  std::vector<double> work((size_t)nchunk * nkx, 0);
  double total = 0;
  for (int ch = 0; ch < nchunk; ch++)
    for (int kx = 0; kx < nkx; kx++) {
      const int r0 = K.lex_row0[kx], r1 = kx + 1 < nkx ? K.lex_row0[kx + 1] : K.nrows;
      for (int r = r0; r < r1; r++) {
        const int nb = K.lex_meta[2 * r] - ch * bpc;
        if (nb <= 0) continue;
        kym[(size_t)ch * nkx + kx] = r - r0;
        work[(size_t)ch * nkx + kx] += std::min(nb, bpc) + 2.0;
      }
      total += work[(size_t)ch * nkx + kx];
    }
  // the kernel walks each chunk's rows as a ky-prefix per kx: check it
  for (int ch = 0; ch < nchunk; ch++)
    for (int kx = 0; kx < nkx; kx++) {
      const int v = kym[(size_t)ch * nkx + kx], r0 = K.lex_row0[kx];
      for (int ky = 0; ky <= v; ky++)
        if (K.lex_meta[2 * (r0 + ky)] <= ch * bpc)
          return fail(MDX_E_ARG, "k-space: chunk rows are not a prefix in ky");
    }
  std::vector<int> pt;
  const double target = total / parts;
  for (int ch = 0; ch < nchunk; ch++) {
    int kx0 = 0;
    double w = 0;
    for (int kx = 0; kx < nkx; kx++) {
      w += work[(size_t)ch * nkx + kx];
      const bool last = kx + 1 == nkx;
      if (w > 0 && (w >= target || last)) {
        pt.insert(pt.end(), {ch, kx0, kx + 1, 0});
        kx0 = kx + 1;
        w = 0;
      }
    }
  }
  // drop the tails of kx without rows in the chunk (kym < 0): kernels skip
  // nothing inside a part, so trim each part to rows that exist
  ptrim.clear();
  for (size_t i = 0; i < pt.size(); i += 4) {
    int ch = pt[i], a = pt[i + 1], b = pt[i + 2];
    while (b > a && kym[(size_t)ch * nkx + b - 1] < 0) b--;
    while (a < b && kym[(size_t)ch * nkx + a] < 0) a++;
    for (int kx = a; kx < b; kx++)
      if (kym[(size_t)ch * nkx + kx] < 0) return fail(MDX_E_ARG, "k-space: chunk rows are not a prefix in kx");
    if (b > a) ptrim.insert(ptrim.end(), {ch, a, b, 0});
  }
  nchunk_out = nchunk;
  return 0;
}

int kspace_chunks(Ctx &c, int kzc, int parts) {
  std::vector<int> kym, ptrim;
  int nchunk = 0, rc;
  if ((rc = kchunks_build(c.K, kzc, parts, kym, ptrim, nchunk))) return rc;
  const int np = (int)ptrim.size() / 4;
  if ((rc = c.d_chunk_kymax.alloc(kym.size())) || (rc = c.d_kparts.alloc(ptrim.size())) ||
      (rc = c.kpf.alloc((size_t)np * c.nq * 3)))
    return rc;
  MDX_HIP(hipMemcpy(c.d_chunk_kymax.p, kym.data(), sizeof(int) * kym.size(), hipMemcpyHostToDevice));
  MDX_HIP(hipMemcpy(c.d_kparts.p, ptrim.data(), sizeof(int) * ptrim.size(), hipMemcpyHostToDevice));
  c.kzc = kzc;
  c.nchunk = nchunk;
  c.kparts = parts;
  c.nkpart = np;
  c.gen++;                             // launch shapes and buffers change
  return 0;
}

int kspace_pass(Ctx &c, int measure, hipStream_t s) {
  return c.mode == MDX_FP64 ? run_kspace<double>(c, measure, s) : run_kspace<float>(c, measure, s);
}

int kspace_add(Ctx &c, hipStream_t s) { return add_kspace(c, s); }

}  // namespace mdx
