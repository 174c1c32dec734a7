"""Host-side mirror of cook's set-up of the force path.

Builds the `mdx_system_t` that cook's globals would hand to forces():
site types, charges, exclusions, site-site tables (initss, cook/simdef.c:
1494-1587), r-space splines (initrspace -> initerfc, cook/simdef.c:
1705-1721), Ewald alpha/kappa (Ewaldtest -> Ewaldparm, sim/ewald.c:1027)
and the linked-cell grid (lcsetup, cook/linklist.c:304).

The numerical set-up is done by a `provider` object exposing the five
helpers of include/mdx.h (setss, erfc_table, ewald_autoset,
ewald_kvectors, lcsetup).  The engine library is the default provider.

Units: cook program units (sim/units.h).  The force field is given in
cook's ble-file units (EvdW in kcal/mol, RvdW in Angstrom, charges in e,
masses in g/mol) and converted like readblend() (cook/simdef.c:620-631,
715-760).
"""
import ctypes as C

import numpy as np

from . import abi

ELECTRON = 408.779826922048          # e in program units (sim/units.h:108)
KCAL = 4184.0
EUNIT = 8.31446261815324             # J/mol per program energy unit
MUNIT = 0.831446261815324            # g/mol per program mass unit
PUNIT = 13806490.0                   # Pa per program pressure unit
EUNIT_KCAL = KCAL / EUNIT            # eunit default (cook/simdef.c:43)


class Provider:
    """Set-up helpers of a library with the mdx.h signatures."""

    def __init__(self, lib, prefix):
        self.lib = lib
        self.p = prefix

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def setss(self, Ei, Ri, Ej, Rj, comb_rule, C2):
        ss = abi.LJ()
        corr = C.c_double(0)
        rc = self._f("setss")(C.byref(ss), Ei, Ri, Ej, Rj, comb_rule, C2, C.byref(corr))
        if rc:
            raise RuntimeError("setss failed (%d)" % rc)
        return ss, corr.value

    def setss_exp6(self, Ei, Ri, ai, Ej, Rj, aj, comb_rule, C2):
        ss = abi.Exp6()
        corr = C.c_double(0)
        rc = self._f("setss_exp6")(C.byref(ss), Ei, Ri, ai, Ej, Rj, aj, comb_rule, C2,
                                   C.byref(corr))
        if rc:
            raise RuntimeError("setss_exp6 failed (%d)" % rc)
        return ss, corr.value

    def setss_buck(self, Ai, Bi, Ci, Aj, Bj, Cj, comb_rule, C2):
        ss = abi.Exp6()
        corr = C.c_double(0)
        rc = self._f("setss_buck")(C.byref(ss), Ai, Bi, Ci, Aj, Bj, Cj, comb_rule, C2,
                                   C.byref(corr))
        if rc:
            raise RuntimeError("setss_buck failed (%d)" % rc)
        return ss, corr.value

    def setss_busing(self, Ei, Ri, Ci, Ej, Rj, Cj, comb_rule, C2):
        ss = abi.Exp6()
        corr = C.c_double(0)
        rc = self._f("setss_busing")(C.byref(ss), Ei, Ri, Ci, Ej, Rj, Cj, comb_rule, C2,
                                     C.byref(corr))
        if rc:
            raise RuntimeError("setss_busing failed (%d)" % rc)
        return ss, corr.value

    def erfc_table(self, grid, maxr, cutoff, alpha, shift):
        nmax = int((maxr * alpha) ** 2 * grid) + 2
        tab = (abi.Erfc * nmax)()
        n = C.c_int(0)
        sgrid = C.c_double(0)
        rc = self._f("erfc_table")(tab, C.byref(n), C.byref(sgrid), grid, maxr,
                                   cutoff, alpha, shift)
        if rc:
            raise RuntimeError("erfc_table failed (%d)" % rc)
        return tab, n.value, sgrid.value

    def ewald_autoset(self, qmax, sumq, cutoff, L, epsr, epsk, kappa0=0.2):
        a = C.c_double(0)
        k = C.c_double(kappa0)
        rc = self._f("ewald_autoset")(C.byref(a), C.byref(k), qmax, sumq, cutoff,
                                      (C.c_double * 3)(*L), epsr, epsk)
        if rc:
            raise RuntimeError("ewald_autoset failed (%d)" % rc)
        return a.value, k.value

    def ewald_kvectors(self, kappa, L):
        k = C.c_double(kappa)
        nk = C.c_int(0)
        rc = self._f("ewald_kvectors")(C.byref(k), (C.c_double * 3)(*L), C.byref(nk))
        if rc:
            raise RuntimeError("ewald_kvectors failed (%d)" % rc)
        return k.value, nk.value

    def lcsetup(self, ns, L, ncell=(0, 0, 0), percell=0.0):
        nc = (C.c_int * 3)(*ncell)
        self._f("lcsetup")(nc, ns, percell, (C.c_double * 3)(*L))
        return list(nc)


def engine_provider():
    return Provider(abi.load_engine(), "mdx_")


class BuiltSystem:
    """An mdx_system_t plus the numpy arrays it points into."""

    def __init__(self):
        self.keep = []
        self.struct = abi.System()
        self.info = {}

    def ptr(self, arr, ctype):
        arr = np.ascontiguousarray(arr)
        self.keep.append(arr)
        return arr.ctypes.data_as(C.POINTER(ctype))


def build_system(ff, counts, L, cutoff, epsr=0.05, epsk=0.05, grid=256,
                 rplus=2.5, rshift=3, epsinf=3e33, ncell=(0, 0, 0),
                 alpha=None, kappa=None, lj_cutoff=None, provider=None,
                 omega=0.6, exp6_evdw_scale=1.0, el_corr=0, el_bg=0, el_diag=0, shell=None):
    """ff: force-field dict as stored in tests/golden/*.json['forcefield']:
         types, type_mass, type_EvdW_kcal, type_RvdW,
         species[{site_type, site_charge_e, site_excl, bonds}]
       counts: molecules per species (cook's N[i]); molecules are ordered
       by species like cook's molec[] (sim/siminit.c:464-647).
       ff["form"] = "exp6" selects the exp-6 form (sim/exp6) with the
       stiffness a of each type in ff["type_parm"]; exp6_evdw_scale
       multiplies its EvdW after the kcal conversion (cook's exp-6
       initcombrule() converts kcal to K once more, sim/exp6/sitesite.c:
       52-54, which system_from_cook reproduces).  ff["form"] = "buck" is
       cook's Buckingham form (sim/buck, POW = 0) in the exp-6 table: A_ii
       [kcal/mol] in type_EvdW_kcal, B_ii [1/A] in type_RvdW, C_ii
       [kcal/mol A^6] in type_parm, ff["comb_rule"] as in sim/buck
       (default 5: geometric A and B).  ff["form"] = "busing" is cook's
       Busing form (sim/busing, POW = 0), also in the exp-6 table: f rho_i
       [kcal/mol] in type_EvdW_kcal (negative: the site interacts with
       nothing), R_i [A] in type_RvdW, C_i [sqrt(kcal/mol) A^3] in
       type_parm (sim/busing/sitesite.c:5-16), comb_rule 0.
       el_corr, el_bg, el_diag: cook's el.corr (Yeh-Berkowitz slab
       correction, bits x=1 y=2 z=4, +8 energy and pressure only), el.bg
       (its charged-background force) and el.diag (1: Ediag in the
       energy), sim/ewald.c:477-492, 865-997.
       shell: POLAR&1 shell-core repulsion, {type name: (kappa, rep)} with
       cook's ATOM columns kappa (divided by the electron charge unit as
       cook/simdef.c:692 does) and rep (nonzero: the type repels shells);
       needs Drudes (ff["polar"]) and the Buckingham or Busing form."""
    prov = provider or engine_provider()
    bs = BuiltSystem()
    S = bs.struct
    nt = len(ff["types"])
    form = ff.get("form", "lj")
    if form not in ("lj", "exp6", "buck", "busing"):
        raise ValueError("site-site form %r: lj, exp6, buck or busing" % form)
    buck = form == "buck"
    busing = form == "busing"
    exp6 = form in ("exp6", "buck", "busing")   # buck and busing fill the exp-6 table
    EvdW = [e * EUNIT_KCAL * (exp6_evdw_scale if exp6 else 1.0) for e in ff["type_EvdW_kcal"]]
    RvdW = list(ff["type_RvdW"])
    C2 = cutoff if lj_cutoff is None else lj_cutoff
    sstab = ((abi.Exp6 if exp6 else abi.LJ) * (nt * nt))()
    corr = np.zeros((nt, nt))
    # combining rule of a pair (sim/lj/sitesite.c:33-58, 66-112): the
    # file's comb_rule, with the square-root energy rule for sites of
    # alpha <= 0, or-ed over the pair; the Kirkwood rule (bit 1 clear) is
    # not restated
    rule = [1] * nt
    if "comb_rule" in ff:
        t_alpha = ff.get("type_alpha", [0.0] * nt)
        rule = [int(ff["comb_rule"]) | (1 if t_alpha[i] <= 0 else 0) for i in range(nt)]
    for i in range(nt):
        for j in range(i + 1):
            if buck:
                # A in the EvdW column, B in RvdW, C in type_parm (kcal/mol,
                # 1/A, kcal/mol A^6; sim/buck/sitesite.c:75-81); one rule
                cr = int(ff.get("comb_rule", 5))
                Cp = [c * EUNIT_KCAL for c in ff["type_parm"]]
                Ab = [e * EUNIT_KCAL for e in ff["type_EvdW_kcal"]]
                ss, c = prov.setss_buck(Ab[i], RvdW[i], Cp[i], Ab[j], RvdW[j], Cp[j], cr, C2)
            elif busing:
                # f rho_i [K] and C_i [sqrt(K) A^3] after the eunit powers
                # of Epow (1 and 1/2, sim/busing/sitesite.c:5-16)
                Eb = [e * EUNIT_KCAL for e in ff["type_EvdW_kcal"]]
                Cb = [c * EUNIT_KCAL ** 0.5 for c in ff["type_parm"]]
                ss, c = prov.setss_busing(Eb[i], RvdW[i], Cb[i], Eb[j], RvdW[j], Cb[j],
                                          int(ff.get("comb_rule", 0)), C2)
            elif EvdW[i] == 0 or EvdW[j] == 0:
                ss, c = (abi.Exp6 if exp6 else abi.LJ)(), 0.0
            else:
                cr = rule[i] | rule[j]
                if not cr & 1:
                    raise ValueError("types %d-%d: the Kirkwood combining rule is not supported"
                                     % (i, j))
                if exp6:
                    a = ff["type_parm"]
                    ss, c = prov.setss_exp6(EvdW[i], RvdW[i], a[i], EvdW[j], RvdW[j], a[j], cr, C2)
                else:
                    ss, c = prov.setss(EvdW[i], RvdW[i], EvdW[j], RvdW[j], cr, C2)
            sstab[i * nt + j] = ss
            sstab[j * nt + i] = ss
            corr[i, j] = corr[j, i] = c
    # nbfixes (sim/setss.c:190-199): the pair's own Emin and RvdW in place of
    # the combining rule; the identical-site case of the same helper, whose
    # rule gives Emin and 2 (RvdW/2) = RvdW back
    for a_, b_, E_, R_ in ff.get("nbfix", []):
        if form != "lj":
            raise ValueError("nbfixes: Lennard-Jones form only")
        i, j = ff["types"].index(a_), ff["types"].index(b_)
        Ek = E_ * EUNIT_KCAL
        ss, c = (abi.LJ(), 0.0) if Ek == 0 else prov.setss(Ek, R_ / 2, Ek, R_ / 2, 1, C2)
        sstab[i * nt + j] = ss
        sstab[j * nt + i] = ss
        corr[i, j] = corr[j, i] = c
    bs.keep.append(sstab)

    polar = ff.get("polar") or {}
    site_type, charge, mass, excl, mol_first = [], [], [], [], [0]
    chargepol, alpha_qq = [], []
    cons_first, cons_ij, cons_bondq = [0], [], []
    for isp, n in enumerate(counts):
        sp = ff["species"][isp]
        ns = len(sp["site_type"])
        masks = [0] * ns
        for i, ex in enumerate(sp["site_excl"]):
            for j in ex:
                masks[i] |= 1 << j
                masks[j] |= 1 << i
        for m in range(n):
            base = mol_first[-1]
            for i in range(ns):
                t = sp["site_type"][i]
                site_type.append(t)
                qtot = sp["site_charge_e"][i] * ELECTRON
                # POLAR: centre charge = total - Drude (cook/simdef.c:715-735)
                a_pol, qd_e = polar.get(ff["types"][t], (0.0, 0.0))
                qd = qd_e * ELECTRON
                chargepol.append(qd)
                alpha_qq.append(a_pol / (qd * qd) if qd else 0.0)
                charge.append(qtot - qd)
                mass.append(ff["type_mass"][sp["site_type"][i]] / MUNIT)
                excl.append(masks[i])
            for (a, b, bl) in sp["bonds"]:
                cons_ij += [base + a, base + b]
                cons_bondq.append(bl * bl)
            mol_first.append(base + ns)
            cons_first.append(len(cons_bondq))
    ns_tot = mol_first[-1]
    nmol = len(mol_first) - 1
    q = np.array(charge)
    qpol = np.array(chargepol)
    qt = q + qpol                     # total site charges (Ewaldparm)
    qmax = np.abs(qt).max() if ns_tot else 0.0
    sumq = float(np.sum(qt * qt))
    if alpha is None or kappa is None:
        alpha, kappa = prov.ewald_autoset(qmax, sumq, cutoff, L, epsr, epsk)
    kappa, nk = prov.ewald_kvectors(kappa, L)
    tab, nerf, sgrid = prov.erfc_table(grid, cutoff + rplus, cutoff, alpha, rshift)
    bs.keep.append(tab)
    if not any(ncell):
        ncell = prov.lcsetup(ns_tot, L)

    S.ns = ns_tot
    S.nmol = nmol
    S.ntypes = nt
    S.mol_first = bs.ptr(np.array(mol_first, np.int32), C.c_int)
    S.site_type = bs.ptr(np.array(site_type, np.int32), C.c_int)
    S.charge = bs.ptr(q, C.c_double)
    S.excl = bs.ptr(np.array(excl, np.uint32), C.c_uint)
    if exp6:
        S.sstab_exp6 = C.cast(sstab, C.POINTER(abi.Exp6))
    else:
        S.sstab = C.cast(sstab, C.POINTER(abi.LJ))
    S.L = (C.c_double * 3)(*L)
    S.cutoff = cutoff
    S.alpha = alpha
    S.kappa = kappa
    S.epsinf = epsinf
    S.alphar = alpha
    S.erfc_sgrid = sgrid
    S.erfc_n = nerf
    S.erfc_tab = C.cast(tab, C.POINTER(abi.Erfc))
    S.ncell = (C.c_int * 3)(*ncell)
    if polar:
        S.chargepol = bs.ptr(qpol, C.c_double)
        S.alpha_qq = bs.ptr(np.array(alpha_qq), C.c_double)
    S.mass = bs.ptr(np.array(mass), C.c_double)
    S.ncons = len(cons_bondq)
    S.cons_first = bs.ptr(np.array(cons_first, np.int32), C.c_int)
    S.cons_ij = bs.ptr(np.array(cons_ij if cons_ij else [0, 0], np.int32), C.c_int)
    S.cons_bondq = bs.ptr(np.array(cons_bondq if cons_bondq else [1.0]), C.c_double)
    S.mol_omega = bs.ptr(np.full(max(nmol, 1), omega), C.c_double)
    S.el_corr = int(el_corr)
    S.el_bg = int(el_bg)
    S.el_diag = int(el_diag)
    if shell:
        if not polar or not exp6:
            raise ValueError("shell-core terms need Drudes and the Buckingham or Busing form")
        qk = np.zeros(nt)
        rep = np.zeros(nt, np.int32)
        for name, (kap, rp) in shell.items():
            t = ff["types"].index(name)
            qk[t] = kap / ELECTRON
            rep[t] = 1 if rp else 0
        S.shell_qkappa = bs.ptr(qk, C.c_double)
        S.shell_rep = bs.ptr(rep, C.c_int)
    # No.f: 3 per site minus constraints (cook subtracts 3 for the total
    # momentum only when it is conserved; callers can override)
    bs.info = dict(alpha=alpha, kappa=kappa, nk=nk, ncell=list(ncell),
                   erfc_n=nerf, erfc_sgrid=sgrid, qmax=qmax, sumq=sumq,
                   ns=ns_tot, nmol=nmol, ncons=len(cons_bondq),
                   corr=corr, sstab=sstab, form=form, erfc_tab=tab,
                   mass=np.array(mass), charge=q, chargepol=qpol,
                   alpha_qq=np.array(alpha_qq), polar=bool(polar),
                   site_type=np.array(site_type, np.int32),
                   mol_first=np.array(mol_first, np.int32),
                   excl=np.array(excl, np.uint32),
                   cons_first=np.array(cons_first, np.int32),
                   cons_ij=np.array(cons_ij, np.int32).reshape(-1, 2),
                   cons_bondq=np.array(cons_bondq))
    return bs
