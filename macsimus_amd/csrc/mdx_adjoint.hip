// Parameter updates and the adjoint of the LJ and exp-6 pair terms, for
// fitting the force field to forces or energies.
//
// The LJ force of a pair (cook/ljqqnp.c:147-206 with the smooth cutoff of
// sim/setss.c:137-306) depends on the constants of its type pair:
//   inner r^2 < C1q:  f/r = E48 (x^7 - x^4/2),  u = E4 (x^6 - x^3),  x = Sq/r^2
//   [C1q, C2q):       f/r = A4 t,              u = A t^2,           t = r^2 - C2q
// (E48 = 2 E24; the kernels use E48 for forces and E4, E24 for energies).
// For weights w_i (dL/dF_i of a loss of the forces) the adjoint is
//   dL/dc = sum_i w_i . dF_i/dc = sum over ordered pairs (i,j) of
//           w_i . (r_i - r_j) df_ij/dc,
// and dE/dc = sum over unordered pairs du_ij/dc.  The kernel recomputes the
// pair terms (nothing is stored from the force call), walks the same cell
// stencil as the pair kernel, one site per lane and one partner type per
// launch row, and keeps the sums in FP64.  C1q moves along the parameters
// such that f stays continuous at C1, so its derivative has no boundary
// term.  The chain to sigma, epsilon through setss() is the caller's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "mdx_internal.h"

using namespace mdx;

// Scratch of mdx_adjoint_lj_noalloc: the staged weights, the per-site
// derivative rows and the summed result, in a fixed array of the code
// object instead of buffers allocated per call.  Each device that loads
// the library holds its own copy until the process ends; host code uses
// only the device address that hipGetSymbolAddress gives for the current
// device, never the array's name.
constexpr size_t kAdjScratchBytes = (size_t)MDX_ADJ_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kAdjScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace mdx {

// the device address of the scratch array on device dev (current), looked
// up once per device
int adjoint_scratch_setup(int dev, double **ptr) {
  if (dev < 0 || dev >= kMaxDevices) return fail(MDX_E_ARG, "adjoint scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MDX_E_GPU, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  if (ptr) *ptr = (double *)g_scratch_ptr[dev];
  return 0;
}

}  // namespace mdx

namespace {

constexpr int kBlock = 256;
constexpr double kPiA = 3.14159265358979323846;

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

// derivatives kept per site and partner type: LJ 4 force + 4 energy,
// exp-6 5 force + 5 energy
template <bool E6>
constexpr int ng_of() { return E6 ? 10 : 8; }

// gsite[(tj * ns + s) * NG + k], s in cell order, for the constants c of
// the pair (ti, tj):
//   LJ:    k = 0..3: sum_j w_s . d  df/dc for c = Sq, E48, A4, C2q
//          k = 4..7: 1/2 sum_j du/dc       for c = Sq, E4, A, C2q
//   exp-6: k = 0..4: sum_j w_s . d  df/dc for c = tau, ef, ru, A4, C2q
//          k = 5..9: 1/2 sum_j du/dc       for c = tau, eu, ru, A, C2q
//   (exp-6 inner: f/r = ef y/r - 6 ru/r^8, u = eu y - ru/r^6, y = e^(tau r))
template <bool E6>
__global__ __launch_bounds__(kBlock) void k_adjoint(
    int ns, int nx, int ny, int nz, double Lx, double Ly, double Lz, int w, double cq,
    const double4 *__restrict__ spos, const int4 *__restrict__ sinfo,
    const int *__restrict__ scell, const int *__restrict__ cstart, int nt,
    const double *__restrict__ lj, const double *__restrict__ wgt, double *__restrict__ gsite) {
  constexpr int NG = ng_of<E6>();
  const int s = blockIdx.x * kBlock + threadIdx.x;
  const int tj_want = blockIdx.y;
  if (s >= ns) return;
  double g[NG];
  for (int k = 0; k < NG; k++) g[k] = 0;
  const double4 pi = spos[s];
  const int4 ii = sinfo[s];
  const int ti = ii.x & 0xffff;
  const double *ss = lj + ((size_t)ti * nt + tj_want) * 8;
  const double C1q = ss[4], C2q = ss[5], A = ss[6], A4 = ss[7];
  const double wx = wgt[3 * ii.w], wy = wgt[3 * ii.w + 1], wz = wgt[3 * ii.w + 2];
  if (C2q > 0) {
    const int c = scell[s];
    const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
    for (int dx = -w; dx <= w; dx++) {
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
            const int4 jj = sinfo[j];
            if ((jj.x & 0xffff) != tj_want || j == s) continue;
            const double4 pj = spos[j];
            const double ddx = pi.x - (pj.x + sx), ddy = pi.y - (pj.y + sy), ddz = pi.z - (pj.z + sz);
            const double rr = ddx * ddx + ddy * ddy + ddz * ddz;
            if (rr >= cq || rr >= C2q) continue;
            if (jj.y == ii.y && ((ii.z >> (jj.x >> 16)) & 1)) continue;  // EXCL
            const double wd = wx * ddx + wy * ddy + wz * ddz;
            if (rr < C1q) {
              if (E6) {
                const double x = sqrt(rr), y = exp(ss[0] * x), i6 = 1 / (rr * rr * rr);
                g[0] += wd * ss[2] * y;
                g[1] += wd * y / x;
                g[2] += -wd * 6 * i6 / rr;
                g[5] += 0.5 * ss[1] * x * y;
                g[6] += 0.5 * y;
                g[7] += -0.5 * i6;
              } else {
                const double Sq = ss[0], E4 = ss[1], E48 = ss[3];
                const double x = Sq / rr, x3 = x * x * x, x4 = x3 * x, x6 = x3 * x3;
                g[0] += wd * E48 * (7 * x6 - 2 * x3) / rr;
                g[1] += wd * (x6 * x - 0.5 * x4);
                g[4] += 0.5 * E4 * (6 * x6 / x - 3 * x * x) / rr;
                g[5] += 0.5 * (x6 - x3);
              }
            } else {
              const double t = rr - C2q;
              constexpr int kf = E6 ? 3 : 2, ke = E6 ? 8 : 6;
              g[kf] += wd * t;
              g[kf + 1] += -wd * A4;
              g[ke] += 0.5 * t * t;
              g[ke + 1] += -A * t;
            }
          }
        }
      }
    }
  }
  double *o = gsite + ((size_t)tj_want * ns + s) * NG;
  for (int k = 0; k < NG; k++) o[k] = g[k];
}

// out[(ti * nt + tj) * ng + k] = sum over sites s of type ti of gsite[tj][s][k],
// one workgroup per (ti, tj, k), fixed order
__global__ void k_adjoint_sum(int ns, int nt, int ng, const int4 *__restrict__ sinfo,
                              const double *__restrict__ gsite, double *__restrict__ out) {
  __shared__ double red[kBlock];
  const int id = blockIdx.x, k = id % ng, tj = (id / ng) % nt, ti = id / (ng * nt);
  double sum = 0;
  for (int s = threadIdx.x; s < ns; s += kBlock)
    if ((sinfo[s].x & 0xffff) == ti) sum += gsite[((size_t)tj * ns + s) * ng + k];
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[id] = red[0];
}

}  // namespace

extern "C" int mdx_set_params(mdx_ctx_t *h, const mdx_lj_t *sstab, const double *charge,
                              const double *chargepol) {
  if (!h) return fail(MDX_E_ARG, "mdx_set_params: NULL handle");
  Ctx &c = h->c;
  MDX_HIP(hipSetDevice(c.device));
  MDX_HIP(hipStreamSynchronize(c.stream));
  if (sstab) {
    if (c.exp6) return fail(MDX_E_ARG, "mdx_set_params: the context has the exp-6 form, not LJ");
    const int nt2 = c.ntypes * c.ntypes;
    std::vector<double> lj = lj_table(sstab, nt2);
    std::vector<float> lj32(lj.begin(), lj.end());
    MDX_HIP(hipMemcpy(c.lj64.p, lj.data(), sizeof(double) * lj.size(), hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(c.lj32.p, lj32.data(), sizeof(float) * lj32.size(), hipMemcpyHostToDevice));
  }
  if (charge || chargepol) {
    if (chargepol && !c.polar) return fail(MDX_E_ARG, "mdx_set_params: system has no Drudes");
    std::vector<double> q(c.ns), qp(c.ns, 0.0);
    MDX_HIP(hipMemcpy(q.data(), c.charge.p, sizeof(double) * c.ns, hipMemcpyDeviceToHost));
    if (c.polar)
      MDX_HIP(hipMemcpy(qp.data(), c.chargepol.p, sizeof(double) * c.ns, hipMemcpyDeviceToHost));
    // the charge list (and the k-space buffers sized by it) stays: a site
    // keeps a charge iff it had one
    for (int i = 0; i < c.ns; i++) {
      if (charge && ((charge[i] != 0) != (q[i] != 0)))
        return fail(MDX_E_ARG, "mdx_set_params: the set of charged sites must not change");
      if (chargepol && ((chargepol[i] != 0) != (qp[i] != 0)))
        return fail(MDX_E_ARG, "mdx_set_params: the set of Drude sites must not change");
    }
    if (charge) std::copy(charge, charge + c.ns, q.begin());
    if (chargepol) std::copy(chargepol, chargepol + c.ns, qp.begin());
    // the charge sums and what depends on them: the background energy
    // of a net charge and Ediag (cook's refusal of a charged system
    // without el.bg is made at mdx_create only, so that a fit may pass
    // through charged states)
    c.sumq2 = 0;
    c.qnet = 0;
    c.qmax = 0;
    for (int i = 0; i < c.ns; i++) {
      c.sumq2 += (q[i] + qp[i]) * (q[i] + qp[i]);
      c.qnet += q[i] + qp[i];
      c.qmax = std::max(c.qmax, std::max(std::fabs(q[i]), std::fabs(qp[i])));
    }
    c.bgE = -kPiA / 2 * (c.qnet / c.alpha) * (c.qnet / c.alpha);
    c.Ediag = c.nq ? ewald_ediag(c.L, c.kappa, c.alpha, c.sumq2) : 0;
    MDX_HIP(hipMemcpy(c.charge.p, q.data(), sizeof(double) * c.ns, hipMemcpyHostToDevice));
    if (c.polar)
      MDX_HIP(hipMemcpy(c.chargepol.p, qp.data(), sizeof(double) * c.ns, hipMemcpyHostToDevice));
  }
  return 0;
}

namespace {

// doubles of scratch the adjoint needs: staged weights [3 ns], derivative
// rows [nt][ns][ng], result [nt^2][ng]
size_t adjoint_scratch_doubles(const Ctx &c, int ng) {
  const size_t nt = c.ntypes;
  return 3 * (size_t)c.ns + nt * c.ns * ng + nt * nt * ng;
}

// the adjoint of the context's site-site form at positions r for weights
// w: res[(a*nt+b)*ng + k] as gsite summed over the sites of type a.  The
// scratch is allocated for the call, or (fixed != NULL) is the caller's
// device array of at least adjoint_scratch_doubles; every entry of it
// that is read is written first in the call.
template <bool E6>
int adjoint_run(Ctx &c, const double *r, const double *w, std::vector<double> &res,
                double *fixed = nullptr) {
  constexpr int ng = ng_of<E6>();
  MDX_HIP(hipSetDevice(c.device));
  const int nt = c.ntypes, nt2 = nt * nt;
  const size_t nrow = c.polar ? 2 * (size_t)c.ns : c.ns;
  DBuf<double> wb, gb, ob;
  double *wd, *gs, *out;
  int rc = 0;
  if (fixed) {
    wd = fixed;
    gs = wd + 3 * (size_t)c.ns;
    out = gs + (size_t)nt * c.ns * ng;
  } else {
    if ((rc = wb.alloc(3 * (size_t)c.ns)) || (rc = gb.alloc((size_t)nt * c.ns * ng)) ||
        (rc = ob.alloc((size_t)nt2 * ng))) {
      wb.release(); gb.release(); ob.release();
      return rc;
    }
    wd = wb.p; gs = gb.p; out = ob.p;
  }
  c.state_on_device = false;           // the device positions are now r
  hipError_t e = hipMemcpyAsync(c.r.p, r, sizeof(double) * 3 * nrow, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(wd, w, sizeof(double) * 3 * c.ns, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess && (rc = bin_device(c)) == 0) {
    k_adjoint<E6><<<dim3(nblk(c.ns), nt), kBlock, 0, c.stream>>>(
        c.ns, c.nc[0], c.nc[1], c.nc[2], c.L[0], c.L[1], c.L[2], c.wst, c.cutoff * c.cutoff,
        (const double4 *)c.spos.p, (const int4 *)c.sinfo.p, c.scan_tmp.p, c.cstart.p, nt, c.lj64.p,
        wd, gs);
    k_adjoint_sum<<<nt2 * ng, kBlock, 0, c.stream>>>(c.ns, nt, ng, (const int4 *)c.sinfo.p, gs,
                                                     out);
    e = hipGetLastError();
  }
  res.assign((size_t)nt2 * ng, 0.0);
  if (!rc && e == hipSuccess)
    e = hipMemcpyAsync(res.data(), out, sizeof(double) * res.size(), hipMemcpyDeviceToHost, c.stream);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(c.stream);
  wb.release();
  gb.release();
  ob.release();
  if (rc) return rc;
  if (e != hipSuccess) return fail(MDX_E_GPU, "adjoint: HIP error");
  return 0;
}

}  // namespace

extern "C" int mdx_set_params_exp6(mdx_ctx_t *h, const mdx_exp6_t *sstab) {
  if (!h || !sstab) return fail(MDX_E_ARG, "mdx_set_params_exp6: NULL argument");
  Ctx &c = h->c;
  if (!c.exp6) return fail(MDX_E_ARG, "mdx_set_params_exp6: the context has the LJ form");
  MDX_HIP(hipSetDevice(c.device));
  MDX_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> t6 = exp6_table(sstab, c.ntypes * c.ntypes);
  std::vector<float> t32(t6.begin(), t6.end());
  MDX_HIP(hipMemcpy(c.lj64.p, t6.data(), sizeof(double) * t6.size(), hipMemcpyHostToDevice));
  MDX_HIP(hipMemcpy(c.lj32.p, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mdx_adjoint_lj(mdx_ctx_t *h, const double *r, const double *w, double *gF,
                              double *gE) {
  if (!h || !r || !w || !gF || !gE) return fail(MDX_E_ARG, "mdx_adjoint_lj: NULL argument");
  Ctx &c = h->c;
  if (c.exp6) return fail(MDX_E_ARG, "mdx_adjoint_lj: the context has the exp-6 form, not LJ");
  std::vector<double> res;
  int rc = adjoint_run<false>(c, r, w, res);
  if (rc) return rc;
  const int nt2 = c.ntypes * c.ntypes;
  for (int p = 0; p < nt2; p++)
    for (int k = 0; k < 4; k++) {
      gF[p * 4 + k] = res[p * 8 + k];
      gE[p * 4 + k] = res[p * 8 + 4 + k];
    }
  return 0;
}

// mdx_adjoint_lj with its scratch in the static array (no device memory
// is allocated or freed): the device address comes from the per-device
// look-up made at mdx_create; sizes beyond the array are refused before
// any device work
extern "C" int mdx_adjoint_lj_noalloc(mdx_ctx_t *h, const double *r, const double *w,
                                      double *gF, double *gE) {
  if (!h || !r || !w || !gF || !gE) return fail(MDX_E_ARG, "mdx_adjoint_lj_noalloc: NULL argument");
  Ctx &c = h->c;
  if (c.exp6) return fail(MDX_E_ARG, "mdx_adjoint_lj_noalloc: the context has the exp-6 form, not LJ");
  const size_t need = adjoint_scratch_doubles(c, ng_of<false>()) * sizeof(double);
  if (need > kAdjScratchBytes)
    return fail(MDX_E_ARG, "mdx_adjoint_lj_noalloc: needs " + std::to_string(need) +
                               " bytes of scratch (3 ns + 8 ntypes ns + 8 ntypes^2 doubles), more "
                               "than MDX_ADJ_SCRATCH_BYTES = " + std::to_string(kAdjScratchBytes));
  MDX_HIP(hipSetDevice(c.device));
  double *ptr = nullptr;
  int rc = adjoint_scratch_setup(c.device, &ptr);
  if (rc) return rc;
  std::vector<double> res;
  if ((rc = adjoint_run<false>(c, r, w, res, ptr))) return rc;
  const int nt2 = c.ntypes * c.ntypes;
  for (int p = 0; p < nt2; p++)
    for (int k = 0; k < 4; k++) {
      gF[p * 4 + k] = res[p * 8 + k];
      gE[p * 4 + k] = res[p * 8 + 4 + k];
    }
  return 0;
}

extern "C" int mdx_adjoint_exp6(mdx_ctx_t *h, const double *r, const double *w, double *gF,
                                double *gE) {
  if (!h || !r || !w || !gF || !gE) return fail(MDX_E_ARG, "mdx_adjoint_exp6: NULL argument");
  Ctx &c = h->c;
  if (!c.exp6) return fail(MDX_E_ARG, "mdx_adjoint_exp6: the context has the LJ form");
  std::vector<double> res;
  int rc = adjoint_run<true>(c, r, w, res);
  if (rc) return rc;
  const int nt2 = c.ntypes * c.ntypes;
  for (int p = 0; p < nt2; p++)
    for (int k = 0; k < 5; k++) {
      gF[p * 5 + k] = res[p * 10 + k];
      gE[p * 5 + k] = res[p * 10 + 5 + k];
    }
  return 0;
}
