// Internal declarations of the batched-replica engine (mdx_batch.hip and
// its k-space, mdx_bkspace.hip).
#pragma once
#include <vector>

#include "mdx_internal.h"

namespace mdx {
namespace batch {

constexpr int kWS = 64;      // sites of a system are padded to whole waves

// per-system parameters read by the kernels
struct DSys {
  double L, iL, cq, sgrid, alpha, cdip;
  int nc, w, cell_off, ncell, lj_off, nt, erf_off, pad_off, ns;
  int srt_off;               // first row of its sites in cell order
  int ks;                    // its k-space descriptor, -1: no charges
};

// per charged system: where its k-space work lives in the batch arrays
struct KSys {
  double L, iL, alpha, cdip;
  int sys;                   // batch system index
  int nq, q_off;             // its charges in the batch charge list
  int kxmax, KXT, KYT, KZT;  // pass-1 table lengths (kx, ky, kz padded to 8)
  int nseg, seg_off;         // its pass-1 segments (64 rows x 8 kz) in the batch list
  int nsplit, jper;          // pass-1 charge splits
  long kp_off;               // its pass-1 partials, in segments
  int chunk_off, lexr_off;   // pass-2 chunk table and lexicographic row starts
  int nkpart;                // pass-2 parts
  long kpf_off;              // its pass-2 partial force rows
  int grps_off;              // its row-group starts per kx (matrix-core pass 2)
  int kcb0, kcb1;            // its coefficient-kernel workgroups
};

// k-space of the batch: the single-system T formulation laid out per
// system, every kernel launched once for all charged systems
struct BK {
  int nks = 0, nqt = 0, nseg = 0, n1 = 0, n2 = 0, n2m = 0, nkcb = 0, kzc = 16, kparts = 8;
  // MDX_MIXED: pass 2 on the matrix cores (units of 128 charges, units2m)
  // with FP32 operands (1) or split-FP16 operands of 3 parts (2)
  int kmfma = 0;
  int shmem_max = 0;         // pass-1 LDS tables of the largest system
  long ncoef = 0;
  std::vector<KSys> hks;
  DBuf<char> ks;
  DBuf<int> qrow, qks;
  DBuf<double> qpos;                         // [nqt] double4
  DBuf<int> row_kx, row_ky, row_kzmax, row_coef0;
  DBuf<int> seg_blk, seg_row0, seg_rowend, seg_ks;
  DBuf<int> units1, units2, units2m, parts;  // int4 each
  DBuf<int> grp_start, grp_base, grp_nb;     // row groups of all systems
  DBuf<float> coefg;
  // kmfma 2: the groups' split-FP16 coefficient blocks (global slots
  // grp_hb, grp_nb8 blocks of 8 kz each) and per-group scales
  DBuf<int> grp_hb, grp_nb8;
  DBuf<uint16_t> coefh;
  DBuf<float> gscale;
  int ngrp = 0;
  DBuf<int> lex_row0, lex_meta, chunk_kymax;
  DBuf<double> kpart, tsum, kpf, kcpart;
  DBuf<double> coef64;
  DBuf<float> coef32;
  void release();
};

struct Batch {
  int mode = MDX_MIXED, nsys = 0, npad = 0, nreal = 0, ncell = 0, device = 0;
  bool exp6 = false;                     // site-site form of every system
  hipStream_t stream = nullptr;
  std::vector<DSys> hs;
  std::vector<double> sumq2, kappa;
  // Ewald energy terms of each system: background of a net charge
  // -pi/2 (Q/alpha)^2 / V and el.diag's Ediag (sim/ewald.c:983-997)
  std::vector<double> ebg;
  DBuf<char> dsys;
  DBuf<double> r, f, r1, q, mass;        // padded rows
  DBuf<int> type, sysid;                 // padded rows
  DBuf<int> cellid, cnt, cstart, cursor, sidx, scell, sinfo;
  DBuf<double> spos;
  DBuf<double> lj64, erf64;
  DBuf<float> lj32, erf32;
  DBuf<double> epart, sacc, lam;         // per sorted site, per system
  DBuf<double> stepp;                    // per system h, T, tau, nf, thermostat
  BK k;
  bool state_on_device = false, use_graph = true, graph_ok = false, timing = false;
  Timer tm;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<mdx_step_t> graph_p;
};

// per-system accumulators of the batch (doubles per system in sacc)
enum {
  B_U = 0, B_EEL = 1, B_VIR = 2, B_PV = 3, B_NPAIR = 9, B_NLJ = 10,   // pair kernel (11)
  B_EK = 11, B_PK = 12,                                                // k-space (7)
  B_M = 18,                                                            // dipole (3)
  B_KIN = 21, B_PKIN = 22,                                             // leap-frog (7)
  B_N = 28
};

// k-space layout of the batch for kz chunk kzc of pass 2 and about kparts
// pass-2 parts for the largest system (others in proportion to their
// k-sets); allocates the batch k-space arrays
int bk_setup(Batch &b, int kzc, int kparts);
// k-space forces of all charged systems added to b.f; energy and tensor
// sums (measure) and dipoles into b.sacc
int bk_forces(Batch &b, int measure);

}  // namespace batch
}  // namespace mdx
