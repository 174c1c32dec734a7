#!/usr/bin/env python3
"""Writes the golden force fixtures under tests/golden/ by running the
reference `cook` program (compiled from the untouched MACSIMUS sources by
oracle/Makefile.ref into oracle/_ref/) on hand-written input decks.

Runs only in the build container, where /root/reference exists.  The
fixtures are data: positions, charges, forces, energies and pressures
that cook printed, plus the force-field parameters of the deck, so that
the tests can rebuild the same system without the reference.

What cook prints and where it comes from:
  * `-v128` makes forces() dump r, q, f per site with 12/9 decimals
    (sim/simils.c:2259-2297, called at cook/forces.c:222).
  * `el.test=-2` runs Ewaldtest once on the initial configuration and
    prints Epot, the virial pressure, the trace pressure and the six
    pressure-tensor entries (sim/ewald.c:1320-1347).  Diagonal entries are
    Pvir[i]/(3V)*Punit; off-diagonal entries are raw Pvir[i]*Punit
    (VO() divides the first three components only).
  * alpha, kappa, Q->N and No.cell come from the protocol lines
    "Ewald final values" and "No.cell=" (sim/ewald.c:1212 ff.,
    cook/linklist.c:341).

Usage:  python3 tests/golden/make_golden.py [case ...]
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
COOK = os.path.join(REPO, "oracle", "_ref", "cookewslc")
COOKPOL = os.path.join(REPO, "oracle", "_ref", "cookewslcpol")

KCAL_PER_KJ = 1.0 / 4.184
TWO16 = 2.0 ** (1.0 / 6.0)


def lj_site(name, mass, sigma, eps_kjmol):
    """MACSIMUS LJ table entry: EvdW = -eps [kcal/mol], RvdW = r_min/2."""
    return dict(name=name, mass=mass, EvdW=-eps_kjmol * KCAL_PER_KJ,
                RvdW=sigma * TWO16 / 2.0 if eps_kjmol else 0.5)


# SPC/E (Berendsen 1987) and Smith-Dang ions, sigma [A], eps [kJ/mol]
SITES = {
    "H": lj_site("H", 1.00800, 0.0, 0.0),
    "O": lj_site("O", 15.99940, 3.166, 0.650),
    "NA": lj_site("NA", 22.98977, 2.35, 0.544),
    "CL": lj_site("CL", 35.45300, 4.40, 0.419),
}

HOH = np.deg2rad(109.47)
HH = 2.0 * np.sin(HOH / 2.0)


def spce_species():
    s, c = np.sin(HOH / 2), np.cos(HOH / 2)
    return dict(name="SPCE", sites=[("H", 0.4238, (s, c, 0.0), []),
                                    ("H", 0.4238, (-s, c, 0.0), [0]),
                                    ("O", -0.8476, (0.0, 0.0, 0.0), [0, 1])],
                bonds=[(0, 2, 1.0), (1, 2, 1.0), (0, 1, HH)])


def ion_species(name, q):
    return dict(name=name, sites=[(name, q, (0.0, 0.0, 0.0), [])], bonds=[])


def write_ble(path, species, types, polar=None, nbfix=()):
    """polar: dict type -> (alphapol [A^3], chargepol [e]) for Drude sites;
    nbfix: rows (type, type, Emin [kcal/mol], R [A]) of the nbfixes table."""
    out = []
    out.append("! hand-written force field for golden fixtures (LJ form)")
    out.append("nspec=%d" % len(species))
    out.append("  nsites=%d  nnbfixes=%d  factor14=0.5  distance14=0  comb_rule=1"
               % (len(types), len(nbfix)))
    out.append("  polar=%d  nparms=0;" % (1 if polar else 0))
    out.append("")
    out.append("Lennard-Jones")
    out.append("! i atom  mass  alpha  EvdW[kcal/mol]  RvdW[A]"
               + ("  alphapol chargepol Esat kappa" if polar else ""))
    for i, t in enumerate(types):
        st = SITES[t]
        line = "  %d %-4s %10.5f  0.000  %.12f  %.12f" % (
            i, t, st["mass"], st["EvdW"], st["RvdW"])
        if polar is not None:
            a, qd = polar.get(t, (0.0, 0.0))
            line += "  %.6f  %.6f  0  0" % (a, qd)
        out.append(line)
    out.append("")
    out.append("nbfixes")
    out.append("! atom atom Eref R")
    for a, b, e, r in nbfix:
        out.append("%s %s %.12g %.12g" % (a, b, e, r))
    out.append("")
    for isp, sp in enumerate(species):
        out.append("!!!!!!!!!!")
        out.append("species %s" % sp["name"])
        out.append("!!!!!!!!!!")
        out.append("i=%d  N=0  config=0  water=0" % isp)
        out.append("ns=%d  nc=%d  nangles=0  ndihedrals=0  nimpropers=0  naromatics=0"
                   % (len(sp["sites"]), len(sp["bonds"])))
        out.append("ndependants=0 ;")
        out.append("")
        out.append("sites")
        out.append("! i atom charge x y z # excluded")
        for i, (t, q, x, exc) in enumerate(sp["sites"]):
            out.append("  %d %-4s %.6f  %.9f %.9f %.9f  %d %s" % (
                i, t, q, x[0], x[1], x[2], len(exc), " ".join(map(str, exc))))
        out.append("")
        out.append("bonds")
        out.append("! i atom i atom K r calc Upot")
        for (i, j, b) in sp["bonds"]:
            ti, tj = sp["sites"][i][0], sp["sites"][j][0]
            out.append(" %d %-4s %d %-4s 0 %.12f %.12f 0.0" % (i, ti, j, tj, b, b))
        for tab in ("angles", "dihedrals", "impropers", "aromatics", "dependants"):
            out.append("")
            out.append(tab)
        out.append("")
        out.append("")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def parse_dump(path, polar=False):
    rows = []
    with open(path) as f:
        f.readline()
        L = [float(x) for x in f.readline().split()[:3]]
        for line in f:
            tok = line.split()
            if polar:
                rows.append([float(x) for x in tok[:14]] + [int(tok[14]), int(tok[15])])
            else:
                rows.append([float(x) for x in tok[:7]] + [int(tok[7]), int(tok[8])])
    a = np.array(rows)
    d = dict(L=np.array(L))
    d["r"] = a[:, 0:3]
    d["q"] = a[:, 3]
    if polar:
        d["rpol"] = a[:, 4:7]
        d["qpol"] = a[:, 7]
        d["f"] = a[:, 8:11]
        d["fpol"] = a[:, 11:14]
    else:
        d["f"] = a[:, 4:7]
    d["mol"] = a[:, -2].astype(np.int32)
    d["site"] = a[:, -1].astype(np.int32)
    return d


def parse_prt(path):
    txt = open(path).read()
    out = {}
    m = re.search(r"Ewald final values: cutoff=(\S+) el.alpha=(\S+) el.kappa=(\S+) Q->N=(\d+)", txt)
    if m:
        out["cutoff"] = float(m.group(1))
        out["alpha_printed"] = float(m.group(2))
        out["kappa_printed"] = float(m.group(3))
        out["Nk"] = int(m.group(4))
    m = re.search(r"initializing Ewald: kappa=K/L=(\S+)", txt)
    if m:
        out["kappa_init"] = float(m.group(1))
    m = re.findall(r"No.cell=\((\d+),(\d+),(\d+)\)", txt)
    if m:
        out["ncell"] = [int(x) for x in m[-1]]
    m = re.search(r"reference set TTTTTT\s*\nEpot = (\S+) K\s*\n"
                  r"sqrt\(sum forces\^2\) = (\S+), max force = (\S+)\s*\n"
                  r".*?\nP \(from E by virial theorem\) = (\S+) \[Pa\]\s*\n"
                  r"P \(trace of pressure tensor\) = (\S+) \[Pa\]\s*\n"
                  r"pressure tensor \[Pa\]:\s*\n\s*(\S+) (\S+) (\S+)\s*\n\s*(\S+) (\S+) (\S+)",
                  txt, re.S)
    if m:
        g = [float(x) for x in m.groups()]
        out["Epot"] = g[0]
        out["fnorm"] = g[1]
        out["fmax"] = g[2]
        out["P_virial_Pa"] = g[3]
        out["P_trace_Pa"] = g[4]
        out["Ptens_printed_Pa"] = g[5:11]
    m = re.search(r"En\.corr=(\S+) \[K AA3\]", txt)
    if m:
        out["Ecorr_KAA3"] = float(m.group(1))
    m = re.search(r"POLAR: (\d+) iterations", txt)
    if m:
        out["scf_iterations"] = int(m.group(1))
    return out


def run(cmd, cwd):
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=900)
    if p.returncode != 0:
        sys.stderr.write(p.stdout.decode(errors="replace")[-3000:])
        raise RuntimeError("cook failed: %s" % cmd)
    return p.stdout.decode(errors="replace")


def ff_arrays(species, types, polar=None):
    """Force-field description stored with every fixture."""
    d = dict(types=types,
             type_mass=[SITES[t]["mass"] for t in types],
             type_EvdW_kcal=[SITES[t]["EvdW"] for t in types],
             type_RvdW=[SITES[t]["RvdW"] for t in types],
             species=[])
    for sp in species:
        d["species"].append(dict(
            name=sp["name"],
            site_type=[types.index(s[0]) for s in sp["sites"]],
            site_charge_e=[s[1] for s in sp["sites"]],
            site_xyz=[list(s[2]) for s in sp["sites"]],
            site_excl=[s[3] for s in sp["sites"]],
            bonds=[list(b) for b in sp["bonds"]]))
    if polar:
        d["polar"] = {t: list(v) for t, v in polar.items()}
    return d


def force_case(name, species, types, counts, L, cutoff, epsr, epsk,
               extra_def="", polar=None, scf=None, md_steps=0, ewald=None, nbfix=()):
    """One Ewaldtest force evaluation on cook's own initial configuration.

    md_steps>0 first runs that many MD steps (SHAKE, leap-frog, Berendsen)
    from the lattice start and then evaluates the forces on the result.
    ewald: cook's el.corr, el.bg, el.diag ({"el_corr": 4, ...}), set in the
    .def deck and stored with the fixture."""
    exe = COOKPOL if polar else COOK
    if ewald:
        extra_def = " ".join([extra_def] + ["el.%s=%d" % (k[3:], v) for k, v in ewald.items()])
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types, polar, nbfix)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff",
                  "el.epsr=%g" % epsr, "el.epsk=%g" % epsk,
                  "h=0.002", "noint=%d" % max(md_steps, 1), "T=298",
                  "epsc=1e-10"]
        if scf:
            lines += scf
        if extra_def:
            lines.append(extra_def)
        lines.append(";")
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        if md_steps:
            open(os.path.join(td, "eq.get"), "w").write(
                'init="crystal"\nthermostat="Berendsen"\ntau.T=0.05\nno=1;\n')
            run([exe, "sys", "eq"], td)
            start = 'init="start"\n'
            os.rename(os.path.join(td, "eq.cfg"), os.path.join(td, "t.cfg"))
        else:
            start = 'init="crystal"\n'
        open(os.path.join(td, "t.get"), "w").write(
            start + "el.test=-2\nno=0;\nel.test=-2;\nel.test=0;\n")
        run([exe, "sys", "t", "-v128"], td)
        dump = parse_dump(os.path.join(td, "t.0.dump"), polar=bool(polar))
        meta = parse_prt(os.path.join(td, "t.prt"))
    meta.update(dict(name=name, counts=counts, L=L, epsr=epsr, epsk=epsk,
                     md_steps=md_steps, ewald=ewald or {}, nbfix=[list(x) for x in nbfix],
                     forcefield=ff_arrays(species, types, polar)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **dump)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, {k: meta[k] for k in ("Epot", "Nk", "ncell") if k in meta})


def traj_case(name, species, types, counts, L, cutoff, nsteps, thermostat="none"):
    """MD trajectory of cook's Shake() (SHAKE + leap-frog, h=2 fs): the
    positions of every force call are dumped by -v128 (at most 10 dumps).
    dump 0 is the initial force evaluation, dump k (k>=1) holds r(t_{k-1})
    and the forces of step k."""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "noint=%d" % nsteps, "T=298", "epsc=1e-12", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="%s"\ntau.T=0.1\nno=1;\n' % thermostat)
        run([COOK, "sys", "md", "-v128"], td)
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "md.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "md.%d.dump" % k)))
            k += 1
        meta = parse_prt(os.path.join(td, "md.prt"))
        txt = open(os.path.join(td, "md.prt")).read()
        m = re.findall(r"omegac=(\S+) \(", txt)
        if m:
            meta["omegac_last"] = float(m[-1])
    out = dict(L=dumps[0]["L"], q=dumps[0]["q"], mol=dumps[0]["mol"], site=dumps[0]["site"],
               r=np.stack([d["r"] for d in dumps]), f=np.stack([d["f"] for d in dumps]))
    meta.update(dict(name=name, counts=counts, L=L, epsr=0.05, epsk=0.05, h=0.002,
                     thermostat=thermostat, ndumps=len(dumps),
                     forcefield=ff_arrays(species, types)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps")


def plb_case(name, species, types, counts, L, cutoff, nsteps):
    """cook's playback file (openplayback() and writeplayback() of
    sim/simils.c:1550-1700) of a short SHAKE MD run with -y and dt.plb=h:
    its bytes, and the positions of every force call (-v128)"""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "noint=%d" % nsteps, "T=298", "epsc=1e-12", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="none"\ndt.plb=0.002\nno=1;\n')
        run([COOK, "sys", "md", "-v128", "-y%d" % sum(counts)], td)
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "md.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "md.%d.dump" % k)))
            k += 1
        raw = open(os.path.join(td, "md.plb"), "rb").read()
    out = dict(L=dumps[0]["L"], r=np.stack([d["r"] for d in dumps]),
               plb=np.frombuffer(raw, np.uint8))
    meta = dict(name=name, counts=counts, L=L, cutoff=cutoff, h=0.002, ndumps=len(dumps),
                plb_bytes=len(raw), forcefield=ff_arrays(species, types))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps", len(raw), "plb bytes")


FRICTION_RUNS = [("tr", "BerendsenCM", ""), ("in", "BerendsenIN", ""),
                 ("fr", "frictions", "T_tr_in=1.3\n")]


def friction_case(name, species, types, counts, L, cutoff, nsteps, runs=FRICTION_RUNS):
    """the decoupled friction thermostats of sim/thermo.c:107-135: from the
    configuration of 5 Berendsen steps (cfg0), three runs of nsteps each
    with thermostat="BerendsenCM" (translations), "BerendsenIN" (the
    intramolecular motion) and "frictions" (both, T_tr_in=1.3), tau.T=0.05;
    stored: cfg0 and each run's final configuration.  Other runs (tag,
    thermostat, extra lines) may be given, e.g. the plain Berendsen
    thermostat of sim/thermo.c:141-160."""
    out = {}
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "T=298", "epsc=1e-12", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="Berendsen"\ntau.T=0.1\nnoint=5\nno=1;\n')
        run([COOK, "sys", "md"], td)
        cfg0 = open(os.path.join(td, "md.cfg"), "rb").read()
        out["cfg0"] = np.frombuffer(cfg0, np.uint8)
        for tag, th, extra in runs:
            open(os.path.join(td, "md.cfg"), "wb").write(cfg0)
            open(os.path.join(td, "md.get"), "w").write(
                'init="start"\nthermostat="%s"\n%stau.T=0.05\nnoint=%d\nno=1;\n' % (th, extra, nsteps))
            run([COOK, "sys", "md"], td)
            out["cfg_" + tag] = np.frombuffer(open(os.path.join(td, "md.cfg"), "rb").read(), np.uint8)
        meta = parse_prt(os.path.join(td, "md.prt"))
    meta.update(dict(name=name, counts=counts, L=L, cutoff=cutoff, epsr=0.05, epsk=0.05,
                     h=0.002, T=298.0, tauT=0.05, T_tr_in=1.3, nsteps=nsteps,
                     runs=[r[:2] for r in runs], forcefield=ff_arrays(species, types)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, [(k, len(v)) for k, v in out.items()])


def nose_case(name, species, types, counts, L, cutoff, nsteps, tauT, pred):
    """Nose-Hoover with the time-reversible velocity predictor in cook's
    Shake() (sim/trvpscale.c, sim/verlet.c, sim/thermo.c): run 1 (crystal,
    Berendsen, 5 steps) saves md.cfg; run 2 starts from it (init="start")
    with thermostat="Nose", tau.T and -p<pred>, every force call dumped by
    -v128.  Stored: both .cfg files as bytes (the start state r, r1, logs
    and the final one) and the dumped positions and forces."""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "T=298", "epsc=1e-12", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="Berendsen"\ntau.T=0.1\nnoint=5\nno=1;\n')
        run([COOK, "sys", "md"], td)
        cfg0 = open(os.path.join(td, "md.cfg"), "rb").read()
        open(os.path.join(td, "md.get"), "w").write(
            'init="start"\nthermostat="Nose"\ntau.T=%g\nnoint=%d\nno=1;\n' % (tauT, nsteps))
        run([COOK, "sys", "md", "-v128", "-p%d" % pred], td)
        cfg1 = open(os.path.join(td, "md.cfg"), "rb").read()
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "md.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "md.%d.dump" % k)))
            k += 1
        meta = parse_prt(os.path.join(td, "md.prt"))
    out = dict(L=dumps[0]["L"], r=np.stack([d["r"] for d in dumps]),
               f=np.stack([d["f"] for d in dumps]),
               cfg0=np.frombuffer(cfg0, np.uint8), cfg1=np.frombuffer(cfg1, np.uint8))
    meta.update(dict(name=name, counts=counts, L=L, cutoff=cutoff, epsr=0.05, epsk=0.05,
                     h=0.002, T=298.0, tauT=tauT, pred=pred, nsteps=nsteps, ndumps=len(dumps),
                     forcefield=ff_arrays(species, types)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps", len(cfg0), len(cfg1), "cfg bytes")


def npt_case(name, species, types, counts, L, cutoff, nsteps, tauT, tauP, P_Pa, pred,
             thermostat="MTK"):
    """The isotropic MTK (or Hoover) barostat with the Nose thermostat and
    the velocity predictor in cook's Shake() (sim/trvpscale.c, sim/verlet.c,
    sim/constrd.c:705-762, sim/thermo.c:11-95): run 1 (crystal, Berendsen,
    5 steps) saves md.cfg; run 2 starts from it with thermostat="MTK",
    tau.T, tau.P, P [Pa], rescale="xyz" and -p<pred>, every force call
    dumped by -v128 (positions, forces and the box of each step).  Stored:
    both .cfg files as bytes and the dumps."""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "T=298", "epsc=1e-12", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="Berendsen"\ntau.T=0.1\nnoint=5\nno=1;\n')
        run([COOK, "sys", "md"], td)
        cfg0 = open(os.path.join(td, "md.cfg"), "rb").read()
        open(os.path.join(td, "md.get"), "w").write(
            'init="start"\nthermostat="%s"\ntau.T=%g\ntau.P=%g\nP=%.10g\nrescale="xyz"\n'
            'noint=%d\nno=1;\n' % (thermostat, tauT, tauP, P_Pa, nsteps))
        run([COOK, "sys", "md", "-v128", "-p%d" % pred], td)
        cfg1 = open(os.path.join(td, "md.cfg"), "rb").read()
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "md.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "md.%d.dump" % k)))
            k += 1
        meta = parse_prt(os.path.join(td, "md.prt"))
        txt = open(os.path.join(td, "md.prt")).read()
    out = dict(L=np.stack([d["L"] for d in dumps]), r=np.stack([d["r"] for d in dumps]),
               f=np.stack([d["f"] for d in dumps]),
               cfg0=np.frombuffer(cfg0, np.uint8), cfg1=np.frombuffer(cfg1, np.uint8))
    meta.update(dict(name=name, counts=counts, L=L, cutoff=cutoff, epsr=0.05, epsk=0.05,
                     h=0.002, T=298.0, tauT=tauT, tauP=tauP, P_Pa=P_Pa, pred=pred,
                     thermostat=thermostat, nsteps=nsteps, ndumps=len(dumps),
                     forcefield=ff_arrays(species, types)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps", len(cfg0), len(cfg1), "cfg bytes")
    return txt


def parse_rdf(path):
    """cook's binary .rdf file (saverdf(), sim/simmeas.c:812-825): records
    <int4 size><data><int4 checksum> (gen/varfile.c); the grid, the number
    of site types, then per type pair (i, j<=i) the rdf_t of sim/rdf.h
    without its leading size field: name[2][8], indx[2], ns[2] (int4),
    4 bytes of alignment, npair, grid, V (double), nmeas, nhist (int4),
    hist[nhist] (unsigned int4)."""
    import struct
    data = open(path, "rb").read()
    recs, o = [], 0
    while o + 4 <= len(data):
        (n,) = struct.unpack_from("<i", data, o)
        o += 4
        if n == 0:
            break
        rec = data[o:o + n]
        o += n
        (chk,) = struct.unpack_from("<i", data, o)
        o += 4
        if (0x4A4B + sum(rec)) & 0xffffffff != chk & 0xffffffff:
            raise RuntimeError("%s: checksum" % path)
        recs.append(rec)
    grid = struct.unpack("<d", recs[0])[0]
    nsites = struct.unpack("<i", recs[1])[0]
    pairs = []
    for rec in recs[2:]:
        names = (rec[0:8].split(b"\0")[0].decode(), rec[8:16].split(b"\0")[0].decode())
        indx = struct.unpack_from("<2i", rec, 16)
        ns = struct.unpack_from("<2i", rec, 24)
        npair, grd, V = struct.unpack_from("<3d", rec, 36)
        nmeas, nhist = struct.unpack_from("<2i", rec, 60)
        hist = np.frombuffer(rec, dtype="<u4", count=nhist, offset=68).copy()
        pairs.append(dict(names=names, indx=list(indx), ns=list(ns), npair=npair, grid=grd, V=V,
                          nmeas=nmeas, nhist=nhist, hist=hist))
    return grid, nsites, pairs


def rdf_case(name, species, types, counts, L, cutoff, nsteps, grid):
    """site-site RDF histograms of cook's measure path (cook/ljqqnp.c:
    218-225): a short MD run with rdf.grid, every force call dumped by
    -v128 (dump 0: initial evaluation, dump k: step k) and the binary
    histograms saved by cook at the end, parsed and as the file's bytes."""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "noint=%d" % nsteps, "T=298", "epsc=1e-12", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="none"\nrdf.grid=%g\nrdf.cutoff=%g\nno=1;\n'
            % (grid, cutoff))
        run([COOK, "sys", "md", "-v128"], td)
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "md.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "md.%d.dump" % k)))
            k += 1
        fn = [f for f in os.listdir(td) if f.endswith(".rdf")]
        g, nsites, pairs = parse_rdf(os.path.join(td, fn[0]))
        raw = open(os.path.join(td, fn[0]), "rb").read()
        meta = parse_prt(os.path.join(td, "md.prt"))
    out = dict(L=dumps[0]["L"], q=dumps[0]["q"], mol=dumps[0]["mol"], site=dumps[0]["site"],
               r=np.stack([d["r"] for d in dumps]))
    for p in pairs:
        out["hist_%d_%d" % tuple(p["indx"])] = p["hist"]
    out["rdf_file"] = np.frombuffer(raw, np.uint8)
    meta.update(dict(name=name, counts=counts, L=L, epsr=0.05, epsk=0.05, h=0.002,
                     ndumps=len(dumps), rdf_grid=g, rdf_nsites=nsites, rdf_cutoff=cutoff,
                     rdf_pairs=[{k: v for k, v in p.items() if k != "hist"} for p in pairs],
                     forcefield=ff_arrays(species, types)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps", [(p["indx"], p["nmeas"], int(p["hist"].sum())) for p in pairs])


def polar_case(name, species, types, counts, L, cutoff, polar):
    """Drude (POLAR=0) force evaluations: Ewaldtest iterates forces() +
    selffield() (sim/ewald.c:1247-1257); every forces() call is dumped with
    the centre and Drude positions/forces (sim/simils.c:2285-2292)."""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types, polar)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "noint=1", "T=298", "scf.eps=-3", "scf.omega=0.7", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "t.get"), "w").write(
            'init="crystal"\nel.test=-2\nno=0;\nel.test=-2;\nel.test=0;\n')
        run([COOKPOL, "sys", "t", "-v128"], td)
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "t.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "t.%d.dump" % k), polar=True))
            k += 1
        meta = parse_prt(os.path.join(td, "t.prt"))
        txt = open(os.path.join(td, "t.prt")).read()
    out = dict(L=dumps[0]["L"], q=dumps[0]["q"], qpol=dumps[0]["qpol"], mol=dumps[0]["mol"],
               site=dumps[0]["site"], r=dumps[0]["r"],
               rpol=np.stack([d["rpol"] for d in dumps]), f=np.stack([d["f"] for d in dumps]),
               fpol=np.stack([d["fpol"] for d in dumps]))
    meta.update(dict(name=name, counts=counts, L=L, epsr=0.05, epsk=0.05, ndumps=len(dumps),
                     scf_omega=0.7, forcefield=ff_arrays(species, types, polar)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps", {k: meta.get(k) for k in ("Epot", "Nk", "ncell", "scf_iterations")})


def polar_traj_case(name, species, types, counts, L, cutoff, polar, nsteps):
    """MD of the Drude system with ASPC k=4 (option -p420: ASPC, sim/pred2.c)
    and one selffield() iteration per step (scf.eps=-1), SHAKE + leap-frog,
    NVE; every force call is dumped (centres and Drude offsets)."""
    with tempfile.TemporaryDirectory() as td:
        write_ble(os.path.join(td, "sys.ble"), species, types, polar)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.001", "noint=%d" % nsteps, "T=298", "epsc=1e-12", "scf.eps=-1", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="none"\nno=1;\n')
        run([COOKPOL, "sys", "md", "-v128", "-p420"], td)
        dumps = []
        k = 0
        while os.path.exists(os.path.join(td, "md.%d.dump" % k)):
            dumps.append(parse_dump(os.path.join(td, "md.%d.dump" % k), polar=True))
            k += 1
        meta = parse_prt(os.path.join(td, "md.prt"))
        txt = open(os.path.join(td, "md.prt")).read()
        meta["aspc_lines"] = re.findall(r"@@@@@@@ initializing .*|scf.omega=\S+ \(stability.*|B\[\d\]=\S+", txt)
    out = dict(L=dumps[0]["L"], q=dumps[0]["q"], qpol=dumps[0]["qpol"], mol=dumps[0]["mol"],
               site=dumps[0]["site"], r=np.stack([d["r"] for d in dumps]),
               rpol=np.stack([d["rpol"] for d in dumps]), f=np.stack([d["f"] for d in dumps]),
               fpol=np.stack([d["fpol"] for d in dumps]))
    meta.update(dict(name=name, counts=counts, L=L, epsr=0.05, epsk=0.05, h=0.001,
                     ndumps=len(dumps), forcefield=ff_arrays(species, types, polar)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, len(dumps), "dumps")


def main(argv):
    subprocess.run(["make", "-s", "-f", "oracle/Makefile.ref", "-j8", "all", "polar"],
                   cwd=REPO, check=True)
    spce = spce_species()
    cases = {
        # 1000 SPC/E, L=31.04 A, 9 A cutoff, Ewald tuned to 0.05 K/A
        "spce1000_lattice": lambda: force_case(
            "spce1000_lattice", [spce], ["H", "O"], [1000], 31.04, 9.0, 0.05, 0.05),
        "spce1000_md": lambda: force_case(
            "spce1000_md", [spce], ["H", "O"], [1000], 31.04, 9.0, 0.05, 0.05,
            md_steps=40),
        # small brine: 3 NaCl + 500 SPC/E, L chosen for ~1.03 g/cm3
        "brine506": lambda: force_case(
            "brine506", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 3, 500], 24.86, 9.0, 0.05, 0.05,
            md_steps=20),
        # the same with an Na-O nbfix (Emin 0.2 kcal/mol at 3.0 A)
        "brine506_nbfix": lambda: force_case(
            "brine506_nbfix", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 3, 500], 24.86, 9.0, 0.05, 0.05,
            md_steps=20, nbfix=[("NA", "O", -0.2, 3.0)]),
        # polarizable: Drude on O (1.44 A^3, -8 e) and on Cl (3.5 A^3, -8 e)
        "polar306": lambda: polar_case(
            "polar306", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 3, 300], 21.10, 9.0,
            {"O": (1.44, -8.0), "CL": (3.5, -8.0)}),
        "polar306_aspc": lambda: polar_traj_case(
            "polar306_aspc", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 3, 300], 21.10, 9.0,
            {"O": (1.44, -8.0), "CL": (3.5, -8.0)}, 8),
        # Ewald corrections of Ewald(2) (sim/ewald.c:865-997): a net charge
        # of +1 e (3 Na, 2 Cl, el.bg=1) gets the background term; the slab
        # (Yeh-Berkowitz) correction in z; a charged box with the slab
        # correction in x and z, its background force and the diagonal term
        "brine505_charged": lambda: force_case(
            "brine505_charged", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 2, 500], 24.86, 9.0, 0.05, 0.05, md_steps=10,
            ewald={"el_bg": 1}),
        "brine506_slab": lambda: force_case(
            "brine506_slab", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 3, 500], 24.86, 9.0, 0.05, 0.05, md_steps=10,
            ewald={"el_corr": 4}),
        "brine505_slabbg": lambda: force_case(
            "brine505_slabbg", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 2, 500], 24.86, 9.0, 0.05, 0.05, md_steps=10,
            ewald={"el_corr": 5, "el_bg": 1, "el_diag": 1}),
        # site-site RDF histograms of the small brine over 3 MD steps
        "brine506_rdf": lambda: rdf_case(
            "brine506_rdf", [ion_species("NA", 1.0), ion_species("CL", -1.0), spce],
            ["H", "O", "NA", "CL"], [3, 3, 500], 24.86, 9.0, 3, 20.0),
        # decoupled friction thermostats on 216 SPC/E from a saved configuration
        "spce216_friction": lambda: friction_case(
            "spce216_friction", [spce], ["H", "O"], [216], 18.6206, 8.0, 6),
        # MTK and Hoover isotropic NPT (Nose thermostat, TRVP k=3) on 216
        # SPC/E from a saved configuration at 100 MPa
        "spce216_npt": lambda: npt_case(
            "spce216_npt", [spce], ["H", "O"], [216], 18.6206, 8.0, 8, 0.05, 0.2, 1e8, 3),
        "spce216_hoover": lambda: npt_case(
            "spce216_hoover", [spce], ["H", "O"], [216], 18.6206, 8.0, 8, 0.05, 0.2, 1e8, 3,
            thermostat="Hoover"),
        # Nose-Hoover (TRVP k=3) on 216 SPC/E from a saved configuration
        "spce216_nose": lambda: nose_case(
            "spce216_nose", [spce], ["H", "O"], [216], 18.6206, 8.0, 8, 0.05, 3),
        # playback file of 216 SPC/E over 4 steps
        "spce216_plb": lambda: plb_case(
            "spce216_plb", [spce], ["H", "O"], [216], 18.6206, 8.0, 4),
        # NVE trajectory of 216 SPC/E (SHAKE + leap-frog)
        "spce216_nve": lambda: traj_case(
            "spce216_nve", [spce], ["H", "O"], [216], 18.6206, 8.0, 8),
        # the plain Berendsen thermostat (tau.T = 0.05 ps, T = 298 K), 8 steps
        # from the configuration of 5 Berendsen steps (cook's md.cfg)
        "spce216_ber": lambda: friction_case(
            "spce216_ber", [spce], ["H", "O"], [216], 18.6206, 8.0, 8,
            runs=[("ber", "Berendsen", "")]),
    }
    todo = argv or list(cases)
    for c in todo:
        cases[c]()


if __name__ == "__main__":
    main(sys.argv[1:])
