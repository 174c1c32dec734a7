// Internal declarations shared by the engine's translation units.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mdx.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#endif

namespace mdx {

// records the message of the last failure of this thread, returns code
int fail(int code, const std::string &msg);

// Ediag of Ewald(0) (sim/ewald.c:477-492): the k-space terms between
// kappa and sqrt(kappa^2+alpha^2), times sumq/(2 pi V), for the box L
double ewald_ediag(const double L[3], double kappa, double alpha, double sumq);

// coefficients of the cell-dipole terms (sim/ewald.c:865-931) per
// direction d: the field on a charge is -cm[d] M_d + cb[d] r_d, the
// energy sum_d ce[d]/2 M_d^2; el_corr selects the Yeh-Berkowitz form,
// otherwise the epsinf form; cb is the charged-background term of el_bg
void ewald_dipole_coefs(int el_corr, int el_bg, double epsinf, double qnet, double V,
                        double cm[3], double ce[3], double cb[3]);

// device address of mdx_adjoint_lj_noalloc's static scratch array on
// device dev (the current one), looked up once per device
int adjoint_scratch_setup(int dev, double **ptr);

// the checks cook's set-up makes on these settings (sim/siminit.c:
// 152-160, sim/ewaldprt.c:16-25, sim/main.c:719-721): a charged system
// needs el.bg and a tin-foil epsinf, the slab correction an infinite one
int ewald_check(int el_corr, int el_bg, double epsinf, double qnet, double sumq2,
                const char *who);

#ifdef __HIPCC__

#define MDX_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      return ::mdx::fail(MDX_E_GPU, std::string(#call ": ") +             \
                                        hipGetErrorString(e_));          \
  } while (0)

// a device buffer owned by a context
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    if (count <= n && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
      return fail(MDX_E_GPU, "hipMalloc failed");
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// number of accumulated scalars per block in the energy/virial reductions
enum Acc {
  ACC_U = 0,      // LJ energy
  ACC_EEL,        // r-space Coulomb energy (incl. exclusion corrections)
  ACC_VIR,        // sum rr*f over LJ terms
  ACC_PV0,        // 6 virial tensor components xx yy zz yz zx xy
  ACC_NPAIR = ACC_PV0 + 6,
  ACC_NLJ,
  ACC_EK,         // k-space energy sum_k w e 4 sum T^2
  ACC_PK0,        // 6 k-space tensor sums (before the 2/(4 pi) factor)
  ACC_M0 = ACC_PK0 + 6,  // dipole moment M (3)
  ACC_S0 = ACC_M0 + 3,   // second moments sum q r_a r_b (6: xx yy zz yz zx xy)
  ACC_SUMQ = ACC_S0 + 6, // sum q^2 (constant, kept for checking)
  ACC_N
};

// Kernel timing by family with HIP events recorded on the launching stream
// around each launch; totals are read (after a stream sync) by mdx_get_timing.
enum TimerFamily { T_BIN = 0, T_PAIR, T_EXCL, T_KPASS1, T_KCOEF, T_KPASS2, T_SHAKE, T_NFAM };

struct Timer {
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> rec;
  size_t used = 0;
  hipEvent_t open[T_NFAM] = {};
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  void start(int fam, hipStream_t s) {
    open[fam] = get();
    if (open[fam]) (void)hipEventRecord(open[fam], s);
  }
  void stop(int fam, hipStream_t s) {
    hipEvent_t e = get();
    if (!e || !open[fam]) return;
    (void)hipEventRecord(e, s);
    rec.push_back({fam, {open[fam], e}});
    open[fam] = nullptr;
  }
  // sums the recorded intervals per family (stream must be synchronized)
  void collect(double *ms, int *count) {
    for (int f = 0; f < T_NFAM; f++) { ms[f] = 0; count[f] = 0; }
    for (auto &r : rec) {
      float t = 0;
      if (hipEventElapsedTime(&t, r.second.first, r.second.second) == hipSuccess) {
        ms[r.first] += t;
        count[r.first]++;
      }
    }
    rec.clear();
    used = 0;
  }
  void release() {
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
    rec.clear();
    used = 0;
  }
};

// k-space layout: rows (kx,ky) sorted by kz limit, 8-wide kz blocks
struct KLayout {
  int kxmax = 0, nrows = 0, nblk = 0, nseg = 0, ntrip = 0, ncoef = 0;
  int kzpad = 0;                       // 8*nblk
  std::vector<int> row_kx, row_ky, row_kzmax, row_coef0;
  std::vector<int> seg_blk, seg_row0, seg_rowend;
  // pass 2: rows in lexicographic order (kx, ky); per row its number of
  // 4-kz coefficient blocks and its first coefficient kz slot
  std::vector<int> lex_row0, lex_meta;  // [kxmax+1], [2*nrows]
  int lex_maxblk = 0;
  // groups of 4 rows (kx, 4g..4g+3) for the matrix-core pass 2: per kx the
  // first group, per group the offset of its coefficient block (floats)
  // and its 4-kz blocks (those of its first, longest row)
  std::vector<int> grp_start, grp_base, grp_nb;  // [kxmax+2], [ngrp], [ngrp]
  long grp_floats = 0;
  // the same groups in 8-kz blocks for the split-FP16 pass 2: per group
  // its first block slot and its number of blocks (ceil of 4 grp_nb / 8)
  std::vector<int> grp_hb, grp_nb8;
  long grp_h8 = 0;                     // block slots of all groups
};

struct Ctx {
  int mode = MDX_MIXED;
  int device = 0;
  hipStream_t stream = nullptr;
  // k-space runs on kstream beside binning, pair and exclusion kernels on
  // stream (fork and join by events), unless overlap is off or kernels
  // are being timed; both streams have the default priority
  hipStream_t kstream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;
  // system
  int ns = 0, nmol = 0, ntypes = 0, nq = 0;
  double L[3] = {0, 0, 0};
  double cutoff = 0, alpha = 0, alphar = 0, kappa = 0, epsinf = 3e33;
  double sumq2 = 0;
  double qmax = 0;                     // largest |charge| (centres, Drudes)
  // Ewald corrections (sim/ewald.c:477-492, 865-997): el.corr, el.bg,
  // el.diag, the net charge, the background energy -pi/2 (Q/alpha)^2
  // (divided by V at each call) and Ediag of the box at set-up
  int el_corr = 0, el_bg = 0, el_diag = 0;
  double qnet = 0, bgE = 0, Ediag = 0;
  int cook_ncell[3] = {0, 0, 0};
  // site data
  DBuf<double> r, f;          // [3ns] input positions, output forces
  DBuf<int> type, mol, molfirst, q_site;
  DBuf<double> charge;        // [ns]
  bool polar = false;          // Drude charges present (cook POLAR)
  DBuf<double> chargepol, alpha_qq;  // [ns]
  // POLAR&1 shell-core: per type qkappa (FP64, FP32) and rep flag
  bool shell = false;
  DBuf<double> shk64;
  DBuf<float> shk32;
  DBuf<int> shrep;
  DBuf<double> spol;           // [4ns] Drude offset, qpol (cell order)
  DBuf<double> qpos;           // [4nq] charge positions and charges
  DBuf<double> qtrig;          // [12nq] per direction cos, sin of 2 pi s and 16 pi s
  // matrix-core pass 1: FP32 anchors e^{2 pi i 8m s} per block of 8 [task][nq]
  // (q folded in along x), then the bases e^{2 pi i s} [3][nq]
  DBuf<float2> qanc;
  // charge adjoint: triple records of the k-set with their S/U sums, the
  // per-split partials of those sums, the per-call buffers
  DBuf<double> qadj_tab, qadj_part, qadj_buf;
  DBuf<int> qadj_rows;
  int qadj_ntrip = 0, qadj_nrows = 0, qadj_nsplit = 1, qadj_ksplit = 1;
  DBuf<unsigned> excl;        // [ns]
  bool exp6 = false;           // site-site form: exp-6 table instead of LJ
  DBuf<double> lj64;          // [ntypes^2*8] LJ or exp-6 constants, C1q C2q A A4
  DBuf<float> lj32;
  DBuf<double> erf64;         // [nerf*8] stable spline form
  DBuf<float> erf32;
  int nerf = 0;
  double erf_sgrid = 0;
  // engine cell grid (pair kernel)
  int nc[3] = {0, 0, 0};
  int wst = 2;                 // stencil half width in cells
  // the pair kernel's stencil columns (dx,dy) are split into parts of
  // pcols columns (0: one part) so that the grid has several waves per
  // SIMD; partial force rows [part][nrow][3] are summed in a fixed order
  int pcols = 0;
  DBuf<double> pf;
  // all-pair path (cook/pair.c) instead of the cell stencil; its parts
  // are slices of the neighbour range (pcols of them, 0: automatic)
  bool allpair = false;
  DBuf<int> cellid, cnt, cstart, cursor, sidx;
  DBuf<double> spos;           // [4ns] x,y,z wrapped, q  (cell order)
  // pairs the pair kernel met within rounding of the cutoff: [bcap] of
  // (site a, site b, its own decision), their count; cook's decision
  // replaces them in the census
  DBuf<int> blist, bcount;
  int bcap = 0;
  DBuf<int> sinfo;             // [4ns] type, mol, excl, site
  DBuf<int> scan_tmp;
  // reductions
  DBuf<double> part;           // per-block partial sums [nblocks*ACC_N]
  int part_blocks = 0;
  DBuf<double> acc;            // [ACC_N] final sums
  // k-space
  KLayout K;
  int nsplit = 1;
  DBuf<int> d_row_kx, d_row_ky, d_row_kzmax, d_row_coef0;
  DBuf<int> d_blk_rows, d_blk_seg0;   // per 8-kz block: rows reaching it, first segment
  DBuf<int> d_seg_blk, d_seg_row0, d_seg_rowend;
  // pass 2: kz chunk of the force kernel (8..64, set by the tuner), the
  // number of chunks and per chunk the largest ky of each kx whose row
  // reaches into the chunk (-1: none); rows reaching a chunk form a
  // ky-prefix for each kx since kzmax(kx,ky) falls with ky
  int kzc = 16, nchunk = 0;
  DBuf<int> d_lex_row0, d_lex_meta, d_chunk_kymax;
  // pass 2 is split into parts (chunk, kx range) of about equal work so
  // that the grid has several waves per SIMD; partial forces per part
  // [nkpart][nq][3] are summed in a fixed order
  int kparts = 16, nkpart = 0;
  // MDX_MIXED: pass 2 on the matrix cores (k_kforce_mfma) instead of FP32
  // FMA chains, from the coefficients also written per row group in the
  // operand order of the matrix instruction (coefg, zero padded)
  int kmfma = 0;                // 0: FMA chains; 1, 2: matrix cores, charge tiles of 32 per wave
  int p1nt = 0;                 // MDX_MIXED pass 1: FP32 FMA (0) or matrix cores, 16 or 8 tiles per wave
  DBuf<int> d_grp_start, d_grp_base, d_grp_nb;
  DBuf<float> coefg;
  // products of the matrix-core pass 2: 0 FP32 operands; 3 (6): each FP32
  // operand as the sum of 2 (3) FP16 parts, scaled by a power of two per
  // row group (coefficients) and 2^8 (phases), and the products of parts
  // whose orders add up to at most 1 (2) formed on the FP16 matrix cores
  // with FP32 accumulation (coefh [slot][part][64 lanes][8], gscale)
  int kprod = 0;
  DBuf<uint16_t> coefh;
  DBuf<float> gscale;
  DBuf<int> d_grp_hb, d_grp_nb8;
  DBuf<int> d_kparts;
  DBuf<double> kpf;
  DBuf<double> kpart;          // pass-1 partials [nsplit][nseg][64][64]
  DBuf<double> tsum;           // reduced T sums [nseg][64][64]
  DBuf<float> coef32;          // pass-2 coefficients [ntrip*16]
  DBuf<double> coef64;
  // MD state
  DBuf<double> r1, pwork;
  DBuf<double> mass;
  DBuf<int> cons_first, cons_ij;
  DBuf<double> cons_bondq, mol_omega;
  // SHAKE sweep counts per species (cook's constrit nit) and the species of
  // each molecule; one species unless mdx_set_shake_species gave more
  int shk_nsp = 1;
  std::vector<int> shk_sp_host;
  DBuf<int> shk_sp;
  DBuf<unsigned long long> shk_nit;
  int ncons = 0;
  bool state_on_device = false;
  // reduction buffers of the MD step: SHAKE/kinetic partials and sums,
  // selffield partials and sums, flags [lambda, SHAKE failed (sticky),
  // max SHAKE iterations (sticky), Drude out of range (sticky)]
  DBuf<double> mdpart, mdacc, sfpart, sfacc, mdflag;
  // Drude ASPC history aa[2..k+1] (aa[0], aa[1] are the Drude rows of r, r1)
  DBuf<double> aspc_hist;
  int aspc_k = -1, aspc_kk = 0;
  // Nose-Hoover: nose = [logs, logs(t)-logs(t-h), En.ext, En.dlogs] and the
  // velocity predictor's history [trvp_k][1 + 3 ns] of length trvp_kk
  DBuf<double> nose, trvp_hist;
  int trvp_k = -1, trvp_kk = 0;
  // isotropic NPT (host): logs, logs1, lambda1 = log L(t) - log L(t-h) and
  // the predictor's histories of logs1 and lambda1 (the site rows are in
  // trvp_hist, whose logs slots NPT leaves unused)
  double npt_logs = 0, npt_logs1 = 0, npt_lam1 = 0, npt_hlogs[9] = {}, npt_hlam[9] = {};
  // configuration generation: bumped by every change of buffers, grid or
  // k-space split, so that a captured step graph is rebuilt
  unsigned gen = 0;
  // the unmeasured MD step captured as a hipGraph (on by default)
  bool use_graph = true, graph_ok = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  unsigned graph_gen = 0;
  int graph_kk = -1;
  mdx_step_t graph_p{};
  // kernel timing with HIP events on the context stream (off by default)
  bool timing = false;
  Timer tm;
  // RDF histograms of the measured calls (off: rdf_grid 0): records
  // [rdf_nrec][rdf_nh], measured calls and their summed box volume
  int rdf_grid = 0, rdf_nh = 0, rdf_nrec = 0, rdf_nmeas = 0;
  double rdf_V = 0, rdf_cutoff = 0;
  DBuf<unsigned long long> rdf_hist;
};

// reduction buffers of the MD step and the Drude selffield
int md_setup(Ctx &c);
// binning of the device positions c.r into the cell-ordered arrays
int bin_device(Ctx &c);
// kernel LJ table [nt2][Sq E4 E24 E48 C1q C2q A A4]
std::vector<double> lj_table(const mdx_lj_t *sstab, int nt2);
// kernel exp-6 table [nt2][tau eu ef ru C1q C2q A A4]
std::vector<double> exp6_table(const mdx_exp6_t *sstab, int nt2);
// k-vector rows of Ewald(0) for box L and kappa (host tables only)
void klayout_build(const double L[3], double kappa, KLayout &K);
// pass-2 tables for kz chunk kzc and about `parts` parts (host): per chunk
// the largest ky of each kx reaching it (kym [nchunk][kxmax+1]) and the
// parts {chunk, kx begin, kx end, 0}
int kchunks_build(const KLayout &K, int kzc, int parts, std::vector<int> &kym,
                  std::vector<int> &parts_out, int &nchunk);
// set-up of the k-space layout for the context box and kappa
int kspace_setup(Ctx &c);
// pass-2 tables for kz chunk kzc (8..64, multiple of 8) and about
// `parts` work parts
int kspace_chunks(Ctx &c, int kzc, int parts);
// k-space on stream s up to the per-part forces of the charges (energy
// and tensor sums into the partial rows when measure, dipole into c.acc)
int kspace_pass(Ctx &c, int measure, hipStream_t s);
// adds the k-space forces and the dipole field to c.f on stream s
int kspace_add(Ctx &c, hipStream_t s);

#endif  // __HIPCC__

}  // namespace mdx

#ifdef __HIPCC__
// the opaque handle of include/mdx.h
struct mdx_ctx_s {
  mdx::Ctx c;
};
#endif
