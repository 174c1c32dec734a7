"""Synthetic seeded configurations of the benchmark workloads.

SPC/E water (Berendsen et al. 1987: q_O=-0.8476 e, q_H=+0.4238 e, O-H 1 A,
HOH 109.47 deg, O-O LJ sigma 3.166 A, eps 0.650 kJ/mol) and Smith-Dang
Na+/Cl- (sigma 2.35/4.40 A, eps 0.544/0.419 kJ/mol), Lorentz-Berthelot
combining, in cook's ble-file units (EvdW [kcal/mol] = -eps, RvdW = r_min/2).
Molecules are placed on a cubic lattice in a random (seeded) order with
random orientations, the way cook's init="crystal" does (sim/simils.c:
436-600).
"""
import numpy as np

AVOGADRO = 6.02214076e23
KCAL_PER_KJ = 1 / 4.184
TWO16 = 2 ** (1 / 6)
HOH = np.deg2rad(109.47)
HH = 2 * np.sin(HOH / 2)
M_H, M_O, M_NA, M_CL = 1.008, 15.9994, 22.98977, 35.453


def _type(mass, sigma, eps_kj):
    return (mass, -eps_kj * KCAL_PER_KJ, sigma * TWO16 / 2 if eps_kj else 0.5)


def brine_forcefield(with_ions=True):
    types = ["H", "O"] + (["NA", "CL"] if with_ions else [])
    tab = {"H": _type(M_H, 0, 0), "O": _type(M_O, 3.166, 0.650),
           "NA": _type(M_NA, 2.35, 0.544), "CL": _type(M_CL, 4.40, 0.419)}
    s, c = np.sin(HOH / 2), np.cos(HOH / 2)
    water = dict(name="SPCE", site_type=[0, 0, 1], site_charge_e=[0.4238, 0.4238, -0.8476],
                 site_xyz=[[s, c, 0], [-s, c, 0], [0, 0, 0]], site_excl=[[], [0], [0, 1]],
                 bonds=[[0, 2, 1.0], [1, 2, 1.0], [0, 1, HH]])
    species = []
    if with_ions:
        species.append(dict(name="NA", site_type=[2], site_charge_e=[1.0], site_xyz=[[0, 0, 0]],
                            site_excl=[[]], bonds=[]))
        species.append(dict(name="CL", site_type=[3], site_charge_e=[-1.0], site_xyz=[[0, 0, 0]],
                            site_excl=[[]], bonds=[]))
    species.append(water)
    return dict(types=types, type_mass=[tab[t][0] for t in types],
                type_EvdW_kcal=[tab[t][1] for t in types],
                type_RvdW=[tab[t][2] for t in types], species=species)


# exp-6 stiffness a of each brine site type (H carries no site-site term)
EXP6_A = {"H": 14.0, "O": 14.0, "NA": 13.0, "CL": 13.5}


def brine_forcefield_exp6(with_ions=True):
    """the brine with the exp-6 form (sim/exp6) in place of LJ: the same
    well depths and minimum distances (EvdW, RvdW), stiffness a from
    EXP6_A (a near 14 follows LJ around the minimum)"""
    ff = brine_forcefield(with_ions)
    ff["form"] = "exp6"
    ff["type_parm"] = [EXP6_A[t] for t in ff["types"]]
    return ff


def brine_c2_exp6():
    """the 194,304-site brine of brine_c2 (same configuration) with the
    exp-6 form"""
    wl = brine_c2()
    wl.update(name="brine_194304_exp6", ff=brine_forcefield_exp6(True))
    return wl


def box_length(counts_masses, density_g_cm3):
    mass_g = sum(n * m for n, m in counts_masses) / AVOGADRO
    return (mass_g / density_g_cm3 * 1e24) ** (1 / 3)


def random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def lattice_config(ff, counts, L, nlat, seed):
    """Positions (molecule-major, cook's order: species by species) of
    counts[i] molecules of each species on an nlat^3 lattice, random slot
    order and random orientations."""
    rng = np.random.default_rng(seed)
    nmol = sum(counts)
    if nmol > nlat ** 3:
        raise ValueError("lattice too small")
    slots = rng.permutation(nlat ** 3)[:nmol]
    a = L / nlat
    centres = np.stack([slots // (nlat * nlat), (slots // nlat) % nlat, slots % nlat], 1) * a + a / 2
    pos = []
    m = 0
    for isp, n in enumerate(counts):
        geo = np.array(ff["species"][isp]["site_xyz"], float)
        com = geo.mean(0)
        geo = geo - com
        rot = random_rotations(rng, n) if len(geo) > 1 else np.tile(np.eye(3), (n, 1, 1))
        pos.append((centres[m:m + n, None, :] + np.einsum("nij,sj->nsi", rot, geo)).reshape(-1, 3))
        m += n
    return np.concatenate(pos)


def brine_c2():
    """64,000 SPC/E + 1,152 Na+/Cl- pairs = 194,304 sites, 1.03 g/cm3,
    random placement on a 41^3 lattice (seed 2), 12 A cutoff."""
    ff = brine_forcefield(True)
    counts = [1152, 1152, 64000]
    L = box_length([(1152, M_NA), (1152, M_CL), (64000, 2 * M_H + M_O)], 1.03)
    r = lattice_config(ff, counts, L, 41, seed=2)
    return dict(name="brine_194304", ff=ff, counts=counts, L=L, cutoff=12.0, r=r)


def polar_c3():
    """Polarizable brine: 32,000 water + 600 Na+/Cl- pairs with a Drude
    charge q_D = -8 e on O (polarizability 1.44 A^3) and on Cl (3.5 A^3),
    spring constant q_D^2/alpha; 97,200 sites, 129,800 charges with the
    Drudes; 1.0 g/cm3, random placement on a 33^3 lattice (seed 3), 10 A
    cutoff.  The Drudes start on their centres (offsets 0)."""
    ff = brine_forcefield(True)
    ff["polar"] = {"O": (1.44, -8.0), "CL": (3.5, -8.0)}
    counts = [600, 600, 32000]
    L = box_length([(600, M_NA), (600, M_CL), (32000, 2 * M_H + M_O)], 1.0)
    r = lattice_config(ff, counts, L, 33, seed=3)
    return dict(name="polar_brine_97200", ff=ff, counts=counts, L=L, cutoff=10.0, r=r, polar=True)


def brine_c5():
    """Force-matching box: 7,856 SPC/E + 72 Na+/Cl- pairs (0.5 mol/kg,
    23,712 sites), 1.0 g/cm3, a full 20^3 lattice in random order (seed 5),
    10 A cutoff."""
    ff = brine_forcefield(True)
    counts = [72, 72, 7856]
    L = box_length([(72, M_NA), (72, M_CL), (7856, 2 * M_H + M_O)], 1.0)
    r = lattice_config(ff, counts, L, 20, seed=5)
    return dict(name="brine_23712", ff=ff, counts=counts, L=L, cutoff=10.0, r=r)


def water_c1():
    """1000 SPC/E on a 10^3 lattice, L=31.04 A, 9 A cutoff (seed 1)."""
    ff = brine_forcefield(False)
    counts = [1000]
    r = lattice_config(ff, counts, 31.04, 10, seed=1)
    return dict(name="spce_3000", ff=ff, counts=counts, L=31.04, cutoff=9.0, r=r)


# reduced LJ units of the replica batch: sigma = 1 A, epsilon = 1 K, m = 1
# program unit, so that the LJ time unit is 1 ps, the reduced temperature
# is T in K and a reduced charge q* is q* program units (q* / e in e)
MUNIT_GMOL = 0.831446261815324
EUNIT_KCAL = 4184.0 / 8.31446261815324


def _lj_type(sigma, eps_K):
    return sigma * TWO16 / 2, -eps_K / EUNIT_KCAL


def replica_forcefield(kind):
    """kind "salt": a restricted-primitive-like binary melt, both species
    sigma = 1, epsilon = 1, reduced charges q* = +-2 (Bjerrum length
    4/T* sigma); kind "ljmix": an 80:20
    LJ mixture A (sigma 1, eps 1) + B (sigma 0.88, eps 0.5), no charges
    (Lorentz-Berthelot)."""
    from .system import ELECTRON
    if kind == "salt":
        types, par, q = ["P", "M"], [(1.0, 1.0), (1.0, 1.0)], [2.0 / ELECTRON, -2.0 / ELECTRON]
    else:
        types, par, q = ["A", "B"], [(1.0, 1.0), (0.88, 0.5)], [0.0, 0.0]
    R = [_lj_type(*p)[0] for p in par]
    E = [_lj_type(*p)[1] for p in par]
    species = [dict(name=t, site_type=[i], site_charge_e=[q[i]], site_xyz=[[0, 0, 0]],
                    site_excl=[[]], bonds=[]) for i, t in enumerate(types)]
    return dict(types=types, type_mass=[MUNIT_GMOL] * 2, type_EvdW_kcal=E, type_RvdW=R,
                species=species)


def replicas_c4(nsys=256, seed=4):
    """The replica batch: nsys independent atomic systems, even ones
    binary salt melts, odd ones LJ mixtures; N ~ U[300, 6000] (even for
    the melts), reduced density U[0.3, 0.9], reduced temperature
    U[0.7, 2.0], cubic boxes, cutoff 2.5 sigma, Ewald at 1e-5 (absolute
    force error in epsilon/sigma), random placement on a lattice and
    Maxwell-Boltzmann velocities (zero total momentum), all from one
    seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nsys):
        kind = "salt" if s % 2 == 0 else "ljmix"
        N = int(rng.integers(300, 6001))
        rho = float(rng.uniform(0.3, 0.9))
        T = float(rng.uniform(0.7, 2.0))
        if kind == "salt":
            N += N % 2
            counts = [N // 2, N // 2]
        else:
            nb = int(round(0.2 * N))
            counts = [N - nb, nb]
        ff = replica_forcefield(kind)
        L = (N / rho) ** (1 / 3)
        nlat = int(np.ceil(N ** (1 / 3) - 1e-9))
        r = lattice_config(ff, counts, L, nlat, seed=int(rng.integers(1 << 30)))
        v = rng.normal(size=(N, 3)) * np.sqrt(T)      # m = 1, k_B = 1
        v -= v.mean(0)
        out.append(dict(name="replica%03d" % s, kind=kind, ff=ff, counts=counts, L=L,
                        cutoff=2.5, r=r, v=v, T=T, rho=rho))
    return out


def build_replica(sysd, provider=None):
    """mdx_system_t of one replica (charged: alpha, kappa from the error
    estimate of Ewaldparm at 1e-5; neutral: the same formulas cannot be
    applied, alpha = 3.5/cutoff and a kappa that gives no k-vectors are
    passed, the k-space sum is empty)."""
    from .system import build_system
    L = [sysd["L"]] * 3
    if sysd["kind"] == "salt":
        return build_system(sysd["ff"], sysd["counts"], L, sysd["cutoff"], epsr=1e-5, epsk=1e-5,
                            provider=provider)
    return build_system(sysd["ff"], sysd["counts"], L, sysd["cutoff"], alpha=3.5 / sysd["cutoff"],
                        kappa=0.5 / sysd["L"], provider=provider)


def degrees_of_freedom(ff, counts):
    """No.f: 3 per site minus constraints minus 3 (total momentum)."""
    nf = 0
    for sp, n in zip(ff["species"], counts):
        nf += n * (3 * len(sp["site_type"]) - len(sp["bonds"]))
    return nf - 3
