/*
 * mdx.h -- C ABI of the MI355X force engine for MACSIMUS `cook`.
 *
 * The engine is a drop-in for cook's force call
 *     void forces(ToIntPtr B, ToIntPtr A)            cook/forces.h:1, cook/forces.c:93-224
 * and the per-step work beside it (SHAKE + leap-frog + Berendsen,
 * sim/constrd.c:622-1081; Drude SCF and ASPC, sim/scf.c:12-393, sim/pred2.c).
 *
 * Conventions are cook's own (sim/units.h:30-108): lengths in Angstrom,
 * energies in K (k_B = 1), time in ps, charges in program units
 * (e = 408.779826922048), Coulomb constant 1, forces in K/Angstrom.
 * Site arrays are cook's `rp` layout: double[3] per site, molecule-major
 * (molec[n].ir, sim/siminit.c:497).  The virial tensor order is
 * xx, yy, zz, yz, zx, xy (sim/simglob.h:487) and is NOT divided by V.
 *
 * Every entry point returns 0 on success or a negative MDX_E* code; the
 * message of the last failure on the calling thread is mdx_last_error().
 * Nothing here calls exit() (cook's ERROR() does, gen/ground.h:157).
 * A context is used from one host thread at a time; all GPU work of a
 * context runs on the context's own HIP stream.
 */
#ifndef MDX_H
#define MDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_OK          0
#define MDX_E_ARG      -1   /* invalid argument or inconsistent system */
#define MDX_E_GPU      -2   /* HIP runtime failure */
#define MDX_E_RANGE    -3   /* site out of box / geometry cook would reject */
#define MDX_E_NOCONV   -4   /* SHAKE or SCF did not converge */
#define MDX_E_NODEV    -5   /* no GPU device visible */

/* precision of a context */
#define MDX_FP64   0   /* plain FP64 GPU path: every kernel in double */
#define MDX_MIXED  1   /* FP32 pair and k-space arithmetic, FP64 accumulation,
                          FP64 exclusions, corrections and global sums */

/* Site-site table of one type pair: cook's sitesite_t by value
   (sim/interpot.h:29-57 + pairaux_t of sim/lj/sitesite.h:44-52).
   u(r) = E4*((Sq/r^2)^6-(Sq/r^2)^3)          for r^2 <  C1q
        = A*(r^2-C2q)^2                        for C1q <= r^2 < C2q
   f(r)/r = E48*y*(y-1/2)*x, x=Sq/r^2, y=x^3  resp.  A4*(r^2-C2q)
   (sim/setss.c:137-306; cook/ljqqnp.c:203-213) */
typedef struct mdx_lj_s {
  double Sq, E4, E24, E48;
  double C1q, C2q, A, A4;
} mdx_lj_t;

/* The same for cook built with the exp-6 form (sim/exp6/sitesite.h:59-75,
   constants of initssaux(), sim/exp6/sitesite.c:78-85):
   u(r) = eu*exp(tau*r) - ru/r^6                for r^2 <  C1q
        = A*(r^2-C2q)^2                          for C1q <= r^2 < C2q
   f(r)/r = ef*exp(tau*r)/r - 6*ru/r^8  resp.  A4*(r^2-C2q)
   with tau = -a/sig, eu = 6 eps e^a/(a-6), ef = eu a/sig,
   ru = eps a sig^6/(a-6) (eps = |E_min| at r = sig, a the stiffness). */
typedef struct mdx_exp6_s {
  double tau, eu, ef, ru;
  double C1q, C2q, A, A4;
} mdx_exp6_t;

/* one interval of cook's hyperbolic r-space splines, ertab_t
   (sim/elst.h:62-68): eru(x)=Bu/(Cu+x)+Au, erd(x)=Bd/(Cd+x)+Ad,
   x=r^2, interval index (int)(x*sgrid) (sim/elst.h:122-130) */
typedef struct mdx_erfc_s {
  double Au, Bu, Cu, Ad, Bd, Cd;
} mdx_erfc_t;

/* Everything cook's globals hand to forces(): the `same site arrays'
   plus the tables initss()/initerfc()/Ewald(0) prepare.  All arrays are
   copied by mdx_create(); the caller keeps ownership. */
typedef struct mdx_system_s {
  int ns;                    /* No.s: LJ/charge centres (without Drudes) */
  int nmol;                  /* No.N: molecules */
  int ntypes;                /* number of site types (nsites in cook) */
  const int *mol_first;      /* [nmol+1] first site of molecule n */
  const int *site_type;      /* [ns] si->st */
  const double *charge;      /* [ns] si->charge (p.u.), central charge */
  const unsigned *excl;      /* [ns] bit k: site mol_first[mol]+k is an
                                EXCL partner (1-2, 1-3; exc->type==EXCL) */
  const mdx_lj_t *sstab;     /* [ntypes*ntypes] sstab[i][j] */
  double L[3];               /* box.L */
  double cutoff;             /* box.cutoff: r-space cutoff (Coulomb) */
  /* Ewald, sim/ewald.c:258-1003 */
  double alpha;              /* el.alpha (k-space screening) */
  double kappa;              /* el.kappa = K/L after the almost-integer fix */
  double epsinf;             /* el.epsinf (3e33 = tin foil) */
  /* r-space splines, sim/erfc.c:119-191 */
  double alphar;             /* Erfc.alpha (el.alphar) */
  double erfc_sgrid;         /* Erfc.sgrid (= grid*alphar^2) */
  int erfc_n;                /* number of intervals */
  const mdx_erfc_t *erfc_tab;/* [erfc_n] Erfc.tab, already shifted */
  /* linked cells, cook/linklist.c:304-382 */
  int ncell[3];              /* No.cell; 0 lets the engine choose */
  /* Drude (POLAR), sim/scf.c; NULL chargepol means nonpolar */
  const double *chargepol;   /* [ns] si->chargepol (p.u.), 0 = no Drude */
  const double *alpha_qq;    /* [ns] si->alpha_qq = alphapol/chargepol^2 */
  /* constraint dynamics (SHAKE=1, VERLET=3), sim/constrd.c:622-1081 */
  const double *mass;        /* [ns] si->mass (p.u.); NULL = no dynamics */
  int ncons;                 /* number of constraints */
  const int *cons_first;     /* [nmol+1] first constraint of molecule n */
  const int *cons_ij;        /* [2*ncons] global site indices i,j */
  const double *cons_bondq;  /* [ncons] squared bond lengths */
  const double *mol_omega;   /* [nmol] SHAKE relaxation omega (= omegac/2) */
  /* site-site form: NULL = Lennard-Jones (sstab); else the exp-6 table
     [ntypes*ntypes] (sstab is then not read).  A batch takes one form
     for all its systems. */
  const mdx_exp6_t *sstab_exp6;
  /* Ewald corrections beyond the self and dipole terms (sim/ewald.c:
     477-492, 865-997; sim/ewaldprt.c:13).  A system with a net charge Q
     (Q/sqrt(sum q^2) > 1e-10; cook's el.epsq is 0) gets cook's background
     energy -pi/2 (Q/alpha)^2 / V in En_el (not in Pvir); it needs el_bg
     and a tin-foil epsinf (|epsinf| >= 1e16), else MDX_E_ARG, as cook's
     ERROR (sim/siminit.c:152-160, sim/ewaldprt.c:16-25).
     el_corr: el.corr, the Yeh-Berkowitz slab correction in place of the
       epsinf dipole term: bit 1, 2, 4 = x, y, z; +8 adds only its energy
       and pressure, not its force; needs |epsinf| >= 1e15.
     el_bg: el.bg (0, 1, 2); nonzero allows a net charge and, with
       el_corr, adds the force q 4 pi Q r_k / V of the charged background
       on each charge in the el_corr directions (and its virial),
       Ballenegger et al.'s term of sim/ewald.c:901-931.  cook's el.bg = 2
       neutralizes the charges at set-up (sim/siminit.c:170-199); a caller
       passes the shifted charges.
     el_diag: el.diag; 1 adds Ediag, the systematic k-space cutoff term
       of the energy computed for the box at mdx_create. */
  int el_corr;
  int el_bg;
  int el_diag;
  /* POLAR&1 shell-core repulsion (cook/ljqqpol.c:41-46, 525-535,
     561-648; cook/simdef.c:646-694, 988-992), for systems with Drude
     charges and an exp-6 table (cook builds it for the Buckingham and
     Busing forms).  Per site type: shell_qkappa, cook's ATOM kappa divided
     by the electron charge unit (si->qkappa; 0: none), and shell_rep
     (cook's rep column; nonzero: the type repels shells).  For a pair
     inside the cutoff whose central type a has qkappa != 0 and whose
     partner type b has rep, with urep(r) = eu e^(tau r) the exponential
     part of the (a,b) table and frep = -urep'(r)/r:
       centre-centre (both centres charged): + qkappa_a urep(r_cc), force
         + qkappa_a frep(r_cc) (r_cc < C2);
       Drude a - centre b (both charged): - qkappa_a urep(r_dc), force
         - qkappa_a frep(r_dc) on the Drude (r_dc < C2);
     and the same with the roles of the two sites swapped.  NULL (both)
     for none. */
  const double *shell_qkappa;
  const int *shell_rep;
} mdx_system_t;

/* what forces() leaves in the globals of sim/simglob.h:436-577 */
typedef struct mdx_out_s {
  double En_el;        /* En.el: Ewald k-space + r-space + self + dipole (SET) */
  double En_pot;       /* En.pot += U_LJ (+ bonded intramolecular = 0 here) */
  double En_vir;       /* En.vir += sum rr*f over LJ terms */
  double Pvir[6];      /* En.Pvir[0..5] += virial tensor (not /V) */
  double En_kspace;    /* k-space energy incl. self and dipole terms */
  long long npair;     /* census: site pairs with rr < cutoff^2, cook's
                          (pairs within rounding of the cutoff decided by
                          cook's arithmetic); -1 if more than
                          max(65536, 2 ns) pairs lie that close to it */
  long long nlj;       /* of those, pairs inside C2q with nonzero LJ */
} mdx_out_t;

typedef struct mdx_ctx_s mdx_ctx_t;

const char *mdx_last_error(void);

/* Creates a context holding the system on the current HIP device.
   mode: MDX_FP64 or MDX_MIXED.  Replaces the one-time set-up that cook
   does in lcsetup() (cook/linklist.c:304), Ewald(0) (sim/ewald.c:334-560)
   and the static scratch re-allocation of linklist()/Ewald(). */
int mdx_create(mdx_ctx_t **ctx, const mdx_system_t *sys, int mode);
int mdx_destroy(mdx_ctx_t *ctx);

/* Replaces forces(B,A) (cook/forces.c:93-224).
   r: configuration A->rp;  f: forces B->rp (overwritten, like sdszero(B)
   at cook/forces.c:183), both [ns*3] doubles, or for a Drude system
   (POLAR, chargepol != NULL) [2*ns*3]: the centres, then the Drude
   offsets rpol / Drude forces fpol as laid out by polarrof()
   (sim/simglob.h).  measure!=0 computes energies, virial and pressure
   tensor (LJqqm/ewpass2m.c variants), otherwise only forces (out->npair
   is still filled).  Host pointers.  Overwrites the device positions, so
   a following mdx_md_steps needs mdx_upload_state again. */
int mdx_forces(mdx_ctx_t *ctx, const double *r, double *f, mdx_out_t *out,
               int measure);

/* Box change between calls (box.L); k-vector set stays as in
   sim/ewald.c:570-613 (ekk recomputed, same ktab). */
int mdx_set_box(mdx_ctx_t *ctx, const double L[3]);
/* The context's box (box.L): as created or set, or as the NPT steps of
   mdx_md_steps left it. */
int mdx_get_box(mdx_ctx_t *ctx, double L[3]);

/* A new system in the same context (No.s, molecules, types, tables or box
   changed: cook re-allocates linklist()'s and Ewald()'s scratch on such a
   change, cook/linklist.c:88-125, sim/ewald.c:367-377).  Everything
   mdx_create builds is built again; the precision and the settings made by
   mdx_set_stencil, mdx_set_kchunk, mdx_set_kspace_mfma,
   mdx_set_kspace_pass1, mdx_set_overlap, mdx_set_graph, mdx_set_timing and
   mdx_set_rdf carry over where the new system allows them (the RDF records
   restart from zero); the MD state must be uploaded again.  On failure
   the context keeps the old system. */
int mdx_resize(mdx_ctx_t *ctx, const mdx_system_t *sys);

/* ---- device-resident MD (the captured step) ---------------------------
   State lives on the GPU: r = cfg[0] (r(t)), r1 = cfg[1] (r(t)-r(t-h)). */
typedef struct mdx_step_s {
  double h;            /* time step [ps] */
  double eps;          /* SHAKE relative tolerance (epsc) */
  int maxit;           /* max SHAKE sweeps per molecule */
  int thermostat;      /* 0 = none (NVE), 1 = Berendsen (sim/thermo.c:141-160),
                          2 = Nose-Hoover with the time-reversible velocity
                          predictor (sim/trvpscale.c, sim/verlet.c:22-36,
                          sim/thermo.c:180-200); not with Drudes;
                          11, 12, 13 = the decoupled friction thermostats
                          of molecular translations, of the intramolecular
                          motion, of both (sim/thermo.c:107-135);
                          22, 23 = MTK, Hoover isotropic NPT (below): every
                          step is measured (the barostat needs the pressure) */
  double T;            /* target temperature [K] */
  double tau;          /* thermostat time constant tau.T [ps] */
  double nf;           /* No.f degrees of freedom */
  int measure_every;   /* energies/virial on every n-th step (cook: noint) */
  /* POLAR: always-stable predictor of the Drude offsets (sim/pred2.c) of
     order aspc_k (cook -p option /100), then scf_iter selffield()
     iterations with mixing scf_omega (0 = the ASPC stability value
     (k+2)/(2k+3), sim/pred2.c:75) */
  int aspc_k;
  int scf_iter;
  double scf_omega;
  /* Nose-Hoover: length of the velocity predictor (cook -p option % 10,
     0..9); the predictor grows to it over the first steps (sim/trvpscale.c) */
  int trvp_k;
  /* friction thermostats 11-13: No.f_tr, the translational degrees of
     freedom (No.f_in = nf - nf_tr), and T_tr_in = T_tr/T_in of 13 */
  double nf_tr;
  double T_tr_in;
  /* 22 = MTK, 23 = Hoover isotropic barostat with the Nose thermostat
     (sim/thermo.c:58-99, sim/trvpscale.c:45-70, sim/siminit.c:604-626;
     cook's rescale="xyz": every site scaled, molecules distorted as in
     rescalecfg(), sim/norm.c:909-1016); not with Drudes.  P: the target
     pressure No.P in program units (Pa / Punit, sim/units.h); tau_P:
     tau.P [ps] (>= 0.999e6: no barostat degree of freedom in No.fx);
     virial: the pressure En.Pref that drives the box, 1 = Pevir (virial
     theorem) or 3 (or 0) = Ptr (trace of the pressure tensor),
     sim/constrd.c:256-262; Ecorr: En.corr of mdx_cutcorr added as
     Ecorr/V^2 (cook's corr&2; 0: none).  tau, T, nf and trvp_k as for
     Nose-Hoover; No.f = nf. */
  double P;
  double tau_P;
  double Ecorr;
  int virial;
} mdx_step_t;

typedef struct mdx_step_out_s {
  mdx_out_t frc;       /* force-call outputs of the last measured step (POLAR:
                          incl. the selffield() terms of En_el and Pvir) */
  double En_self;      /* POLAR: En.self, Drude self energy */
  double scf_maxerr;   /* POLAR: max |d rpol| q_D/omega of the last iteration */
  double En_kin;       /* kinetic energy (VERLET=3 leap-frog estimate) */
  double Pkin[6];      /* kinetic tensor (not /V) */
  double virc;         /* constraint virial En.virc */
  double PTvirc[6];    /* constraint virial tensor */
  double lambda;       /* last Berendsen scaling factor */
  int shake_iter_max;  /* max SHAKE sweeps used by a molecule */
  /* Nose-Hoover, last measured step: En.ext (VERLET=3 form), logs after
     the step and En.dlogs (sim/thermo.c:180-200) */
  double En_ext, logs, dlogs;
  double En_kin_tr;    /* kinetic energy of the molecular translations (En.kin_tr/2) */
  /* NPT (22, 23), last step: box after it; En.Pref that drove it; En.V =
     the volume it was measured at; En.lambda[0] = log L(t) and
     En.dlambda[0] (predicted); En_ext there is No.fx T logs(t) + the
     barostat and thermostat kinetic terms + V P (sim/thermo.c:69-84) */
  double L[3];
  double Pref, V;
  double npt_lambda, npt_dlambda;
  /* SHAKE sweeps summed over the molecules with constraints and over the
     steps of the call: cook's constrit[sp].nit[COR_BONDS] summed over the
     species (sim/constrd.c:930); per species by mdx_get_shake_nit */
  long long shake_nit;
} mdx_step_out_t;

/* Replaces one selffield() iteration (sim/scf.c:12-240, POLAR=0):
   r, f: 2*ns rows (centres, then Drude offsets / Drude forces) as left by
   mdx_forces; the Drude offsets in r are moved, the Drude forces are
   passed to the centres in f; out->En_el and out->Pvir get the
   centre-Drude Ewald correction and the Drude virial.  *Eself = En.self,
   *maxerr = largest |d rpol| q_D / omega. */
int mdx_selffield(mdx_ctx_t *ctx, double *r, double *f, mdx_out_t *out,
                  double omega, double *Eself, double *maxerr);

/* State transfer of cfg[0]->rp (r) and cfg[1]->rp (r1), [ns*3] doubles,
   for a Drude system [2*ns*3]: the Drude rows of r are rpol(t) (ASPC
   aa[0]) and those of r1 the ASPC history aa[1] (sim/pred2.c:27-54). */
int mdx_upload_state(mdx_ctx_t *ctx, const double *r, const double *r1);
int mdx_download_state(mdx_ctx_t *ctx, double *r, double *r1);
/* Runs nsteps of Shake() (sim/constrd.c:622) with forces computed on the
   GPU.  Drude systems: each step is the ASPC prediction (sim/pred2.c:54-99,
   the predictor length grows by one per step up to aspc_k), scf_iter
   forces()+selffield() iterations, then the SHAKE step.  Errors:
   MDX_E_NOCONV (SHAKE did not converge), MDX_E_RANGE (a Drude left
   el.minqq), checked at every measured step and at the last one. */
int mdx_md_steps(mdx_ctx_t *ctx, const mdx_step_t *p, int nsteps,
                 mdx_step_out_t *out);
/* Unmeasured steps in the steady state (for Drudes: once the predictor has
   its full length) are replayed from a hipGraph of one step, captured on
   first use and again after any change of step parameters or engine
   configuration; on = 0 launches every step eagerly.  Kernel timing
   (mdx_set_timing) also runs eagerly.  Default: on. */
int mdx_set_graph(mdx_ctx_t *ctx, int on);
/* The k-space passes of a force call run on a second stream of the
   context, beside binning, pair and exclusion kernels, and join before
   their forces are added (fork/join inside a captured step too); on = 0
   runs everything on one stream.  Results are the same bits either way.
   Kernel timing (mdx_set_timing) runs on one stream.  Default: on. */
int mdx_set_overlap(mdx_ctx_t *ctx, int on);

/* ASPC state of a Drude system: predictor length k (0..8), length
   reached kk (0..k) and the older history aa[2..k+1] as k arrays of ns*3
   doubles (hist = NULL: zeros, like the alloczero of sim/pred2.c:52).
   cook keeps kk and aa[2..] in statics of sim/pred2.c:36-38; these calls
   let a caller restart or inspect them.  mdx_md_steps with another
   aspc_k than the context holds restarts the predictor (kk = 0). */
int mdx_set_aspc(mdx_ctx_t *ctx, int k, int kk, const double *hist);
int mdx_get_aspc(mdx_ctx_t *ctx, int *k, int *kk, double *hist);

/* Isotropic NPT state (thermostats 22, 23): logs = cfg[0]->logs, logs1 =
   cfg[1]->logs, lambda1 = cfg[1]->lambda[0] (log L(t) - log L(t-h); the
   box itself is the context's, mdx_set_box), the velocity predictor's
   length k (0..9), the length reached kk (0..k) and its history
   hist [k][2 + 3 ns]: per level logs1, lambda1, then r1 of the sites
   (hist = NULL: zeros).  cook keeps k, kk and the history in statics of
   sim/constrd.c:673-684 (rows of No.eq: logs, lambda[3], rp) and logs,
   lambda in the cfg.  The predictor (k, kk and the site rows) is the one
   Nose-Hoover uses; logs, logs1 and lambda1 are separate from its state. */
int mdx_set_npt(mdx_ctx_t *ctx, double logs, double logs1, double lambda1, int k, int kk,
                const double *hist);
int mdx_get_npt(mdx_ctx_t *ctx, double *logs, double *logs1, double *lambda1, int *k,
                int *kk, double *hist);

/* Nose-Hoover state: logs = cfg[0]->logs, logs1 = cfg[1]->logs
   (logs(t) - logs(t-h)), the velocity predictor's length k (0..9), the
   length reached kk (0..k) and its history hist [k][1 + 3 ns]: per level
   logs1, then r1 of the sites (hist = NULL: zeros).  cook keeps k, kk and
   the history in statics of sim/constrd.c:673-684 and logs in the cfg;
   mdx_md_steps with another trvp_k than the context holds restarts the
   predictor (kk = 0).  Both logs start at 0 in mdx_create. */
int mdx_set_nose(mdx_ctx_t *ctx, double logs, double logs1, int k, int kk, const double *hist);
int mdx_get_nose(mdx_ctx_t *ctx, double *logs, double *logs1, int *k, int *kk, double *hist);

/* RDF histograms of the measured force calls, as cook records them for
   rdf.grid and rdf.cutoff: initrdf() (sim/simmeas.c:655-767), the counts
   of cook/ljqqnp.c:218-225 (for Drude sites, of the centres,
   cook/ljqqpol.c:477-485) and advancerdf() (sim/simmeas.c:777-787).
   mdx_set_rdf starts them from zero: grid > 0 one record per pair of site
   types (i, j<=i) in the order i = 0.., j = 0..i; grid < 0 one record of
   all pairs; grid = 0 off (the default).  cook truncates rdf.grid to an
   integer (abs() of a double); cutoff in [0, 1e4] gives
   nhist = (int)(cutoff*|grid| + 1.0000001) bins (at most 100000).  A pair
   is counted when it is not excluded and closer than the r-space cutoff,
   in bin (int)(r*|grid|) when that is below nhist.  Every measured force
   call (mdx_forces with measure, measured steps of mdx_md_steps) adds its
   pairs, one measurement and the box volume; for Drude systems only the
   last SCF iteration of a step is measured, where cook counts every
   iteration's forces() (the same g(r), counts divided by the iterations).
   mdx_get_rdf: the records x bins counts (hist NULL: only the sizes), the
   measurements and the summed volume (cook's nmeas and V).  cook keeps
   unsigned 32-bit counts; these are 64-bit. */
int mdx_set_rdf(mdx_ctx_t *ctx, int grid, double cutoff);
int mdx_get_rdf(mdx_ctx_t *ctx, unsigned long long *hist, long n, int *nrec, int *nhist,
                int *nmeas, double *V);

/* ---- diagnostics (no counterpart in cook) -----------------------------
   mdx_cook_cells: cook's linked-cell index of each site for the context's
   cook grid (cook/linklist.c:304-337 with No.cell), to compare the cell
   membership.  mdx_set_timing / mdx_get_timing: HIP-event timing of the
   kernel families (bin, pair, excl, kpass1, kcoef, kpass2, shake) on the
   context stream; ms[i], count[i] accumulate since the last get.
   mdx_set_stencil: cell-stencil half width w (1..6) of the pair kernel
   (a wider one is taken where the box needs it) and the split of its
   (2w+1)^2 stencil columns into parts of pcols columns (0: one part),
   which sets the waves in flight.  w = 0 selects the all-pair path of
   cook/pair.c (minimum image per site pair, for small systems and as a
   cross-check) with pcols slices of the neighbour range (0: automatic).
   mdx_set_kchunk: kz chunk (8..64, multiple of 8) of the k-space force
   kernel, which sets its register footprint, and the number of parts
   (1..256) its k-set is split into, which sets the waves in flight.
   mdx_autotune: times reps force calls per candidate on the positions r
   (NULL: the uploaded MD state) with HIP events and keeps the fastest
   stencil w (1..3, binning + pair time) and kz chunk and parts (pass-2
   time) and stencil columns per pair part; choice[4] = {w, kz chunk,
   parts, columns}, ms[2] = their times per call (either may be NULL).
   mdx_info: iv[0..n) = engine cells x,y,z, stencil w, charges, k rows,
   kz blocks, k segments, k triples, k-sum splits, kx max, mode, kz chunk,
   kz chunks, parts asked, parts made, stencil columns per pair part,
   matrix-core pass 2 (0/1/2), pass-1 tiles (0 = FMA), pass-2 products
   (0/3/6) (n may be 20).
   mdx_set_kspace_mfma (MDX_MIXED): k-space pass 2 on the matrix cores
   (FP32 MFMA, FP32 products and sums as the FMA variant) with on = 1 or
   2 tiles of 32 charges per wave (2 share each coefficient operand), or
   as FP32 FMA chains (on = 0); mdx_autotune tries all three.  Default: 1
   in MDX_MIXED.
   mdx_set_kspace_pass1 (MDX_MIXED): k-space pass 1 as FP32 FMA chains
   (tiles = 0), on the matrix cores as units of 8 k rows with all their
   kz per wave (tiles = 1: FP32 MFMA 16x16x4, sums added to FP32 totals
   every 128 charges and to FP64 per charge split), the same units with
   two waves each, which share the row operands and split the kz blocks
   (tiles = 2), on the FP16 matrix cores with split-FP16 operands (the
   products of mdx_set_kspace_products: 3, or 6 when it is 0 or 6) as row
   units (tiles = 4..7: units of 8 rows, 8 per workgroup, 64
   charges per table chunk (4); units of 4 rows, 16 per workgroup, 64
   charges (5, the default where kz stays below 64 and the system has no
   Drudes; with Drudes the default is 2); units of 4 rows, 16
   per workgroup, two table sets of 32 charges (6); units of 8 rows, 8 per
   workgroup, two sets of 32 charges (7)), or on the FP32 matrix cores with
   16 or 8 tiles of 4 k rows and one 8-kz block per wave (FP64 flush every
   32 charges).
   mdx_autotune keeps the fastest.
   mdx_set_kspace_products (MDX_MIXED, with the matrix-core pass 2): np =
   0 forms pass 2's products of FP32 operands on the FP32 matrix cores;
   np = 3 writes each FP32 operand as the sum of two FP16 parts (22 of its
   24 bits) and forms the three products of parts of order <= 1 on the
   FP16 matrix cores with FP32 sums (terms left out: 2^-22 relative);
   np = 6 uses three parts (all 24 bits) and the six products of order <=
   2 (terms left out: 2^-33).  The operands are scaled by powers of two
   (per 4-row group for the coefficients) to stay in FP16's normal range.
   The tuner leaves this choice alone.  Default: 6 in MDX_MIXED. */
int mdx_set_kspace_mfma(mdx_ctx_t *ctx, int on);
int mdx_set_kspace_products(mdx_ctx_t *ctx, int np);
int mdx_set_kspace_pass1(mdx_ctx_t *ctx, int tiles);
int mdx_cook_cells(mdx_ctx_t *ctx, const double *r, int *cellid);
int mdx_set_timing(mdx_ctx_t *ctx, int on);
int mdx_get_timing(mdx_ctx_t *ctx, double *ms, int *count, int n);
int mdx_set_stencil(mdx_ctx_t *ctx, int w, int pcols);
int mdx_set_kchunk(mdx_ctx_t *ctx, int kzc, int parts);
int mdx_autotune(mdx_ctx_t *ctx, const double *r, int reps, int *choice, double *ms);
int mdx_info(mdx_ctx_t *ctx, int *iv, int n);

/* ---- SHAKE relaxation omega per species (cook's omegac optimization) ----
   cook's default omegac < 0 tunes the relaxation of each species during the
   run from the sweep counts (doconstrit(), sim/norm.c:32-78, with cook's own
   random trial steps).  The engine keeps that on the caller's side: it
   counts the sweeps per species and takes the new omega per species.
   mdx_set_shake_species: species index of each molecule (mol_species[nmol],
     values in [0, nspec), nspec <= 64; NULL: one species).
   mdx_set_shake_omega: the relaxation omega (= omegac/2, constrit[sp].omega)
     of each species for the following steps (omega[nspec]; each in (0, 2)).
   mdx_get_shake_nit: the sweeps of the molecules with constraints of each
     species summed over the steps of the last mdx_md_steps call (nit[nspec],
     cook's constrit[sp].nit[COR_BONDS] of those steps).
   mdx_resize goes back to one species and keeps the omega of the new
   system's mol_omega. */
int mdx_set_shake_species(mdx_ctx_t *ctx, const int *mol_species, int nspec);
int mdx_set_shake_omega(mdx_ctx_t *ctx, const double *omega);
int mdx_get_shake_nit(mdx_ctx_t *ctx, long long *nit, int nspec);

/* ---- force-field fitting (no counterpart in cook) ----------------------
   mdx_set_params: new site-site table sstab [ntypes^2] (mdx_setss) and/or
   charges (centre charges [ns], Drude charges [ns]) for the next calls;
   NULL keeps the old ones.  alpha, kappa, the k-vectors and the set of
   charged sites stay; the net charge, its background energy and Ediag
   follow the new charges (cook's refusal of a charged system without
   el_bg is made by mdx_create only).
   mdx_adjoint_lj: adjoint of the LJ terms at positions r (as mdx_forces)
   for the weights w [ns*3] = dL/dF of a loss of the site forces; for each
   type pair (a,b) (a = central site) gF[(a*nt+b)*4 + k] = sum_i w_i .
   dF_i/dc_k with c = (Sq, E48, A4, C2q) and gE[(a*nt+b)*4 + k] = dEpot/dc_k
   with c = (Sq, E4, A, C2q), each unordered pair split half and half
   between (a,b) and (b,a).  The pair terms are recomputed in FP64.
   mdx_set_params_exp6 / mdx_adjoint_exp6: the same for a context of the
   exp-6 form (mdx_set_params then takes charges only), with
   gF[(a*nt+b)*5 + k] for c = (tau, ef, ru, A4, C2q) and
   gE[(a*nt+b)*5 + k] for c = (tau, eu, ru, A, C2q) of mdx_exp6_t. */
int mdx_set_params(mdx_ctx_t *ctx, const mdx_lj_t *sstab, const double *charge,
                   const double *chargepol);
int mdx_adjoint_lj(mdx_ctx_t *ctx, const double *r, const double *w, double *gF,
                   double *gE);
int mdx_set_params_exp6(mdx_ctx_t *ctx, const mdx_exp6_t *sstab);
/* No counterpart in cook.  For callers that cannot allocate device memory
   (code under HIP graph capture, tight force-matching loops, harnesses
   that forbid hipMalloc): mdx_adjoint_lj with the same arguments and
   results, its scratch (staged weights, per-site derivative rows, summed
   result: 3 ns + 8 ntypes ns + 8 ntypes^2 doubles) in a static device
   array of MDX_ADJ_SCRATCH_BYTES per device instead of buffers allocated
   per call.  Larger systems get MDX_E_ARG before any device work.  The
   array is shared by every call on a device and keeps the last call's
   contents; calls that overlap on one device race. */
#define MDX_ADJ_SCRATCH_BYTES (8u << 20)
int mdx_adjoint_lj_noalloc(mdx_ctx_t *ctx, const double *r, const double *w, double *gF,
                           double *gE);
int mdx_adjoint_exp6(mdx_ctx_t *ctx, const double *r, const double *w, double *gF,
                     double *gE);
/* mdx_adjoint_charges: adjoint of the Ewald electrostatics (r-space with
   cook's erfc spline, EXCL corrections, k-space on the force call's
   k-vector set, self, background, Ediag and cell-dipole terms) with respect
   to the site charges, at positions r (as mdx_forces) for the weights w
   [ns*3] = dL/dF of a loss of the site forces: gF[i] = sum_k w_k . dF_k/dq_i
   and gE[i] = dEn_el/dq_i (program units; 0 for uncharged sites).  Every
   force term is bilinear in the charges, so gF and gE equal central
   differences of mdx_forces in q_i up to rounding.  FP64 in either mode;
   systems with Drude charges get MDX_E_ARG.  No counterpart in cook: it
   replaces differences of the force call in charge fitting. */
int mdx_adjoint_charges(mdx_ctx_t *ctx, const double *r, const double *w, double *gF,
                        double *gE);

/* ---- batched replicas ---------------------------------------------------
   Many independent atomic systems (one site per molecule: LJ with cook's
   smooth cutoff, point charges, Ewald; no exclusions, constraints or
   Drudes) of uneven size in one batch: per system what one cook process
   does with forces() (cook/forces.c:93-224) and the leap-frog + Berendsen
   step of Shake() (sim/constrd.c:622-1081 for molecules without
   constraints), each with its own box, cutoff, tables, k-vector set,
   time step and thermostat, every kernel launched once for all systems.
   Cook runs one system per process; these calls have no counterpart
   beyond that per-system equivalence.
   mdx_batch_create: sys[nsys] as for mdx_create (cubic boxes; nmol == ns,
   ncons == 0, chargepol == NULL, masses required; ncell is ignored: each
   system gets cells of at least cutoff/w, w = 1 until mdx_batch_set_stencil
   or mdx_batch_autotune).  mode: as in mdx_create (the k-space passes
   in FP32 with FP64 accumulation in MDX_MIXED).
   Site arrays of the batch calls are the systems' [ns*3] arrays one after
   another (system order).  out[nsys] per system. */
typedef struct mdx_batch_s mdx_batch_t;
int mdx_batch_create(mdx_batch_t **b, int nsys, const mdx_system_t *sys, int mode);
int mdx_batch_destroy(mdx_batch_t *b);
/* forces(B,A) of every system; measure as for mdx_forces */
int mdx_batch_forces(mdx_batch_t *b, const double *r, double *f, mdx_out_t *out,
                     int measure);
/* MD state r, r1 of all systems (as mdx_upload_state per system) */
int mdx_batch_upload_state(mdx_batch_t *b, const double *r, const double *r1);
int mdx_batch_download_state(mdx_batch_t *b, double *r, double *r1);
/* nsteps leap-frog steps of all systems; p[nsys] per system (h, T, tau,
   nf, thermostat are used, thermostat 0 or 1 (Berendsen) only;
   measure_every of p[0] for all); out[nsys]
   filled at the last measured step (frc, En_kin, Pkin, lambda).
   Unmeasured steps are replayed from one hipGraph of the whole batch step
   (mdx_batch_set_graph, default on), re-captured when p changes. */
int mdx_batch_md_steps(mdx_batch_t *b, const mdx_step_t *p, int nsteps,
                       mdx_step_out_t *out);
int mdx_batch_set_graph(mdx_batch_t *b, int on);
/* cell-stencil half width w (1..4) of all systems (cells >= cutoff/w; a
   system whose box is under (2w+1)/w cutoffs takes a wider one, up to 8) */
int mdx_batch_set_stencil(mdx_batch_t *b, int w);
/* k-space work split: kz chunk kzc (8, 16, 24, 32) of the force pass and
   about kparts parts (1..256) for the largest k-set, the others in
   proportion to their k-sets */
int mdx_batch_set_kspace(mdx_batch_t *b, int kzc, int kparts);
/* MDX_MIXED: k-space pass 2 on the matrix cores: on = 1 with FP32
   operands (as mdx_set_kspace_mfma), on = 2 with split-FP16 operands of
   3 parts and 6 products per FP32 product (as mdx_set_kspace_products 6),
   0 FP32 FMA; mdx_batch_autotune tries all three.  Default: 2 in MDX_MIXED. */
int mdx_batch_set_kspace_mfma(mdx_batch_t *b, int on);
/* times the force call on r (HIP events): the stencil w = 1..3, then kz
   chunk x parts; keeps the fastest; choice[3] = {w, kz chunk, parts},
   ms[2] = force-call times after each stage (either may be NULL) */
int mdx_batch_autotune(mdx_batch_t *b, const double *r, int *choice, double *ms);
/* HIP-event timing of the batch kernel families, in the slots of
   mdx_get_timing (bin, pair, -, kpass1, -, kpass2, leap-frog); steps run
   eagerly while timing is on */
int mdx_batch_set_timing(mdx_batch_t *b, int on);
int mdx_batch_get_timing(mdx_batch_t *b, double *ms, int *count, int n);
/* iv[0..n) = systems, sites, padded sites, cells, charged systems, mode,
   w, charges, k segments, pass-1 units, pass-2 units, kz chunk, parts,
   matrix-core pass 2 (0/1) */
int mdx_batch_info(mdx_batch_t *b, int *iv, int n);

/* ---- pressure assembly and cutoff correction (host arithmetic) -------- */

/* measurePconstraints() (sim/constrd.c:115-265) for PRESSURETENSOR=7: all
   in program units (pressure x Punit = Pa; sim/units.h). */
typedef struct mdx_pressure_s {
  double vir;          /* En.vir: LJ virial + constraint virial + En.el
                          (- 2 En.self with Drudes) */
  double Pvir[6];      /* (Pvir + PTvirc)/V, xx yy zz yz zx xy */
  double Pkin[6];      /* Pkin * Pkinq / V */
  double Ptens[6];     /* Pvir + Pkin */
  double Ptr, Ptr_c;   /* trace(Ptens)/3, + Ecorr/V^2 */
  double Pevir, Pevir_c; /* (2 Ekin Pkinq + vir)/(3V), + Ecorr/V^2 */
} mdx_pressure_t;
/* so: outputs of a measured step (a force call alone: frc filled, the
   rest zero); Pkinq: cook's No.Pkinq (1, or No.f0/No.f with corr&4);
   Ecorr: from mdx_cutcorr; polar: the system has Drudes. */
int mdx_pressure(const mdx_step_out_t *so, const double L[3], double Pkinq,
                 double Ecorr, int polar, mdx_pressure_t *p);
/* cutcorr() (sim/siminit.c:665-722): En.corr = sum over intermolecular
   site pairs of corr(type_i,type_j), corr [ntypes*ntypes] from mdx_setss.
   The energy correction is Ecorr/V, the pressure correction Ecorr/V^2. */
int mdx_cutcorr(const mdx_system_t *sys, const double *corr, double *Ecorr);

/* ---- set-up helpers restating cook's init code (product, not oracle) -- */

/* combining rule (sim/lj/sitesite.c:66-112) + smooth cutoff (sim/setss.c:
   137-306).  EvdW in K (negative), RvdW in Angstrom, comb_rule as in the
   ble file, C2 = LJcutoff.  Writes ss; returns the homogeneous cutoff
   correction corr (sim/setss.c:455-501) in *corr if corr != NULL. */
int mdx_setss(mdx_lj_t *ss, double EvdW_i, double RvdW_i, double EvdW_j,
              double RvdW_j, int comb_rule, double C2, double *corr);
/* The same for the exp-6 form: combrule() of sim/exp6/sitesite.c:60-76
   (sig = RvdW_i+RvdW_j, or sqrt(4 RvdW_i RvdW_j) for comb_rule&2;
   eps = sqrt(EvdW_i EvdW_j); a = sqrt(a_i a_j); comb_rule&1 must be set,
   cook has no Kirkwood rule for exp-6) and the smooth cutoff and tail
   correction of setss() on its u and f.  EvdW in K (negative). */
int mdx_setss_exp6(mdx_exp6_t *ss, double EvdW_i, double RvdW_i, double a_i,
                   double EvdW_j, double RvdW_j, double a_j, int comb_rule, double C2,
                   double *corr);
/* The Buckingham form (cook built with sim/buck, POW = 0),
   u = A_ij exp(-B_ij r) - C_ij/r^6, into the exp-6 table (tau = -B_ij,
   eu = A_ij, ef = A_ij B_ij, ru = C_ij; initssaux() of
   sim/buck/sitesite.c:178-188) with the combining rules of combrule()
   (sim/buck/sitesite.c:113-170): comb_rule&3 for B (0 Waldman-Hagler,
   1 geometric, 2 arithmetic, 3 additive 1/B), comb_rule&12 for A (0
   Waldman-Hagler, 4 geometric, 8 Kong, 12 Busing), C geometric.  A in K,
   B in 1/Angstrom, C in K Angstrom^6 (the .ble columns EvdW, RvdW,
   parm[0] after the eunit conversion). */
int mdx_setss_buck(mdx_exp6_t *ss, double A_i, double B_i, double C_i, double A_j,
                   double B_j, double C_j, int comb_rule, double C2, double *corr);
/* The Busing form (cook built with sim/busing, POW = 0),
   u = f rho_ij exp((R_ij - r)/rho_ij) - C_ij/r^6 with f = 0.05 e^2/A^2,
   into the exp-6 table (tau = -1/rho_ij, ef = f exp(R_ij/rho_ij),
   eu = rho_ij ef, ru = C_ij; initssaux() of sim/busing/sitesite.c:87-108)
   with combrule() (sim/busing/sitesite.c:63-80): rho_ij f = EvdW_i +
   EvdW_j (0, the pair switched off, if either is negative), R_ij =
   RvdW_i + RvdW_j, C_ij = C_i C_j; comb_rule must be 0.  EvdW = f rho_i in
   K, RvdW = R_i in Angstrom, C_i in sqrt(K) Angstrom^3. */
int mdx_setss_busing(mdx_exp6_t *ss, double EvdW_i, double RvdW_i, double C_i, double EvdW_j,
                     double RvdW_j, double C_j, int comb_rule, double C2, double *corr);

/* initerfc()+makesplines() with SPLINE=-2, SUBGRID=1 (sim/erfc.c:119-191,
   sim/splines.c:23-158): fills tab[0..*n) (caller provides room for
   mdx_erfc_size(...) entries). */
int mdx_erfc_size(int grid, double maxr, double alpha);
int mdx_erfc_table(mdx_erfc_t *tab, int *n, double *sgrid, int grid,
                   double maxr, double cutoff, double alpha, int shift);

/* Ewaldparm(): alpha, kappa from epsr, epsk (sim/ewald.c:1027-1067);
   qmax = charges.max, sumq = charges.sumq */
int mdx_ewald_autoset(double *alpha, double *kappa, double qmax, double sumq,
                      double cutoff, const double L[3], double epsr,
                      double epsk);
/* Ewald(0): the almost-integer kappa fix and the number of k-vectors
   Q->N (sim/ewald.c:484-522); kappa is updated in place */
int mdx_ewald_kvectors(double *kappa, const double L[3], int *nk);
/* lcsetup(): No.cell from No.s (cook/linklist.c:294-337) */
int mdx_lcsetup(int ncell[3], int ns, double percell, const double L[3]);

#ifdef __cplusplus
}
#endif
#endif /* MDX_H */
