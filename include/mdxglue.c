/* Glue between cook and the MI355X engine (include/mdx.h), compiled into
   cook with -DMDX; the hooks that call it are in cook/forces.c:93,
   sim/scf.c:12, cook/cook.c (mdxinit), sim/main.c:935 and
   sim/mainres.c:423 (RDF).  This file is part of the boundary: it reads
   cook's globals (sim/simglob.h) and is compiled against cook's own
   headers by the test suite; it is not built into libmdx.so. */
#include "ground.h"
#include "simglob.h"
#include "interpot.h"
#include "elst.h"
#include "simdef.h"                       /* nsites: number of site types */
#include "norm.h"                         /* constrit, doconstrit() */
#include "mdx.h"

static mdx_ctx_t *mdx;
static int *mdx_molfirst, *mdx_type, *mdx_consfirst, *mdx_consij;
static unsigned *mdx_excl;
static double *mdx_q, *mdx_qpol, *mdx_aqq, *mdx_mass, *mdx_bondq, *mdx_omega;
static mdx_lj_t *mdx_ss;
static mdx_erfc_t *mdx_erf;

static void mdxcheck(int rc, const char *what) {
  if (rc) ERROR(("%s: %s", what, mdx_last_error()))
}

void mdxinit(int mode) /* after initss(), initrspace(), Ewald(0), lcsetup() */
{
  int n, i, j, s = 0, c = 0, nt = nsites;   /* nsites = # of site types */
  mdx_system_t S = {0};

  allocarray(mdx_molfirst, No.N + 1);
  allocarray(mdx_consfirst, No.N + 1);
  allocarray(mdx_type, No.s);  allocarray(mdx_q, No.s);  allocarray(mdx_qpol, No.s);
  allocarray(mdx_aqq, No.s);   allocarray(mdx_mass, No.s); allocarray(mdx_excl, No.s);
  allocarray(mdx_consij, 2 * No.c); allocarray(mdx_bondq, No.c); allocarray(mdx_omega, No.N);

  loop (n, 0, No.N) {                      /* molecule-major, as A->rp */
    molecule_t *mn = molec + n;
    specinfo_t *sp = spec[mn->sp];
    mdx_molfirst[n] = s;  mdx_consfirst[n] = c;
    mdx_omega[n] = constrit ? constrit[mn->sp].omega : fabs(omegac) / 2;
    loop (i, 0, mn->ns) {
      siteinfo_t *si = sp->si + i;
      exception_t *e;
      mdx_type[s + i] = si->st;
      mdx_q[s + i] = si->charge;           /* centre charge (POLAR: total - Drude) */
#ifdef POLAR
      mdx_qpol[s + i] = si->chargepol;
      mdx_aqq[s + i] = si->alpha_qq;
#endif
      mdx_mass[s + i] = si->mass;
      mdx_excl[s + i] = 0;
      for (e = si->exc; e && e->type != END; e++) {
        if (e->type != EXCL) ERROR(("MDX: only EXCL exceptions are supported"))
        mdx_excl[s + i] |= 1u << e->indx;
      }
    }
    loop (j, 0, mn->nc) {                  /* constraints: si[j].pair, si[j].bondq */
      mdx_consij[2 * c] = s + sp->si[j].pair[0];
      mdx_consij[2 * c + 1] = s + sp->si[j].pair[1];
      mdx_bondq[c++] = sp->si[j].bondq;
    }
    s += mn->ns;
  }
  mdx_molfirst[No.N] = s;  mdx_consfirst[No.N] = c;

  allocarray(mdx_ss, nt * nt);             /* sitesite_t -> Sq E4 E24 E48 C1q C2q A A4 */
  loop (i, 0, nt) loop (j, 0, nt) {
    sitesite_t *ss = &sstab[i][j];
    mdx_lj_t *m = mdx_ss + i * nt + j;
    m->Sq = ss->a.Sq;  m->E4 = ss->a.E4;  m->E24 = ss->a.E24;  m->E48 = ss->a.E48;
    m->C1q = ss->C1q;  m->C2q = ss->C2q;  m->A = ss->A;        m->A4 = ss->A4;
  }
  /* sstab already holds cook's nbfixes (setss(), sim/setss.c:190-199):
     the copy carries them with no further glue */
  /* the spline table can be passed as is (ertab_t has the same six doubles) */
  S.erfc_tab = (const mdx_erfc_t *)Erfc.tab;
  S.erfc_n = mdx_erfc_size(Erfc.grid, Erfc.maxr, el.alpha);
  S.erfc_sgrid = Erfc.sgrid;

  S.ns = No.s;  S.nmol = No.N;  S.ntypes = nt;
  S.mol_first = mdx_molfirst;  S.site_type = mdx_type;  S.charge = mdx_q;
  S.excl = mdx_excl;  S.sstab = mdx_ss;
  VV(S.L, =box.L)  S.cutoff = box.cutoff;
  S.alpha = el.alpha;  S.kappa = el.kappa;  S.epsinf = el.epsinf;  S.alphar = el.alpha;
  VV(S.ncell, =No.cell)
#ifdef POLAR
  S.chargepol = mdx_qpol;  S.alpha_qq = mdx_aqq;
#endif
  S.mass = mdx_mass;  S.ncons = c;  S.cons_first = mdx_consfirst;  S.cons_ij = mdx_consij;
  S.cons_bondq = mdx_bondq;  S.mol_omega = mdx_omega;
  /* Ewald corrections of Ewald(2): el.corr, el.bg, el.diag (sim/ewald.c:477-492, 865-997) */
  S.el_corr = el.corr;  S.el_bg = el.bg;  S.el_diag = el.diag;
#if defined(POLAR) && POLAR&1
  {                                        /* shell-core: si->qkappa, rep of the site types */
    static double *qk;  static int *rep;
    allocarray(qk, nt);  allocarray(rep, nt);
    loop (i, 0, nt) { qk[i] = sitedef[i].kappa;  rep[i] = sitedef[i].rep != 0; }
    S.shell_qkappa = qk;  S.shell_rep = rep;
  }
#endif
  mdxcheck(mdx_create(&mdx, &S, mode), "mdx_create");
  {                                        /* species of the molecules: sweep counts */
    int *sp;
    allocarray(sp, No.N);
    loop (n, 0, No.N) sp[n] = molec[n].sp;
    mdxcheck(mdx_set_shake_species(mdx, sp, nspec), "mdx_set_shake_species");
    free(sp);
  }
}

/* cook's omegac optimization (sim/main.c:1621, doconstrit() of sim/norm.c)
   with the engine's sweep counts, after every step of mdxsteps(.., 1) */
void mdxconstrit(void)
{
  long long nit[64];
  double om[64];
  int sp;
  mdxcheck(mdx_get_shake_nit(mdx, nit, nspec), "mdx_get_shake_nit");
  loop (sp, 0, nspec) constrit[sp].nit[COR_BONDS] = (int)nit[sp];
  doconstrit(omegac);
  loop (sp, 0, nspec) om[sp] = constrit[sp].omega;
  mdxcheck(mdx_set_shake_omega(mdx, om), "mdx_set_shake_omega");
}

/* forces(B,A) body under MDX; cook/forces.c:93 */
void mdxforces(ToIntPtr B, ToIntPtr A)
{
  mdx_out_t o;
  int i;
  mdxcheck(mdx_forces(mdx, (const double *)A->rp, (double *)B->rp, &o, measure),
           "mdx_forces");
  En.el = o.En_el;                          /* set, like Ewald(2) */
  En.pot += o.En_pot;
  En.vir += o.En_vir;
  loop (i, 0, 6) En.Pvir[i] += o.Pvir[i];
}

#ifdef POLAR
/* selffield() body under MDX; sim/scf.c:12 */
int mdxselffield(ToIntPtr B, ToIntPtr A, double omega)
{
  mdx_out_t o = {0};
  double eself, maxerr;
  int i;
  mdxcheck(mdx_selffield(mdx, (double *)A->rp, (double *)B->rp, &o, omega, &eself, &maxerr),
           "mdx_selffield");
  En.el += o.En_el;
  loop (i, 0, 6) En.Pvir[i] += o.Pvir[i];
  En.self = eself;
  return maxerr < scf.eps;
}
#endif

/* RDF histograms of the measured calls (advancerdf() of cook/forces.c:220
   is not called under MDX) */
#include "simmeas.h"
static double mdx_rdfcut;

void mdxrdfinit(double grid, double cutoff)
{
  mdx_rdfcut = cutoff;
  if (grid) mdxcheck(mdx_set_rdf(mdx, (int)grid, cutoff), "mdx_set_rdf");
}

void mdxrdfsave(double grid)
{
  int nrec, nh, nmeas, i, j, k, p = 0;
  double V;
  unsigned long long *h;
  if (!grid) return;
  mdxcheck(mdx_get_rdf(mdx, NULL, 0, &nrec, &nh, NULL, NULL), "mdx_get_rdf");
  allocarray(h, (long)nrec * nh);
  mdxcheck(mdx_get_rdf(mdx, h, (long)nrec * nh, NULL, NULL, &nmeas, &V), "mdx_get_rdf");
  loop (i, 0, grid < 0 ? 1 : nsites) loopto (j, 0, i) {
    rdf_t *r = rdf[i][j];
    if (r) {
      loop (k, 0, nh) r->hist[k] += (unsigned4)h[(long)p * nh + k];
      r->nmeas += nmeas;  r->V += V;
    }
    p++;
  }
  free(h);
  mdxcheck(mdx_set_rdf(mdx, (int)grid, mdx_rdfcut), "mdx_set_rdf");  /* counted once */
}

/* the device-resident steps of Shake() (sim/constrd.c:622): noint steps
   with the state on the GPU, cfg[0]->rp = r(t), cfg[1]->rp = r(t)-r(t-h) */
void mdxsteps(ToIntPtr *cfg, int noint)
{
  mdx_step_t p = { h, epsc, 0, thermostat == T_BERENDSEN, T, tau.T, No.f, noint,
#ifdef POLAR
                   option('p') / 100 % 10, 1, option('^') ? scf.omega : 0
#else
                   0, 0, 0
#endif
                 };
  mdx_step_out_t so;
  int i;
  mdxcheck(mdx_upload_state(mdx, (double *)cfg[0]->rp, (double *)cfg[1]->rp), "upload");
  mdxcheck(mdx_md_steps(mdx, &p, noint, &so), "mdx_md_steps");
  mdxcheck(mdx_download_state(mdx, (double *)cfg[0]->rp, (double *)cfg[1]->rp), "download");
  En.el = so.frc.En_el;  En.pot = so.frc.En_pot;  En.vir = so.frc.En_vir;
  loop (i, 0, 6) En.Pvir[i] = so.frc.Pvir[i];
  En.kin = so.En_kin;  En.virc = so.virc;
}
