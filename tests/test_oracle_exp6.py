"""The exp-6 site-site form (sim/exp6) in the oracle, pinned by what the
mathematics fixes; cook's exp-6 build cannot be compiled here (its
simopt.h is generated by cook's configure step), so no cook run pins it.

- The table of a pair: u(sig) = -eps and u'(sig) = 0 (eps = |E_min| at
  r = sig), for the combining rules of sim/exp6/sitesite.c:60-76.
- The smooth cutoff: u and f/r continuous at C1 and the quartic's own
  consistency A4 = -4 A, which holds only at the fixed point for C1.
- The tail correction against an adaptive quadrature of its two
  integrals.
- The pair sum: energy, virial, forces and tensor of a neutral exp-6 fluid
  against a brute-force sum over all pairs with the minimum image.
- The oracle's and the engine's set-up helpers agree.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.system import EUNIT_KCAL, Provider, build_system

PAIRS = [  # (E_i kcal, R_i, a_i, E_j, R_j, a_j, comb_rule)
    (-0.1553, 1.7766, 14.0, -0.1553, 1.7766, 14.0, 1),
    (-0.1300, 1.3190, 13.0, -0.1001, 2.4694, 13.5, 1),
    (-0.1300, 1.3190, 13.0, -0.1553, 1.7766, 14.0, 3),
]
C2 = 9.0


def providers():
    out = [Provider(oracle.load(), "oracle_")]
    try:
        out.append(Provider(abi.load_engine(), "mdx_"))
    except OSError:
        pass
    return out


def table(prov, p):
    Ei, Ri, ai, Ej, Rj, aj, cr = p
    return prov.setss_exp6(Ei * EUNIT_KCAL, Ri, ai, Ej * EUNIT_KCAL, Rj, aj, cr, C2)


def u_of(t, r):
    return t.eu * math.exp(t.tau * r) - t.ru / r ** 6


def fr_of(t, r):
    """force / r"""
    return t.ef * math.exp(t.tau * r) / r - 6 * t.ru / r ** 8


@pytest.mark.parametrize("p", PAIRS)
def test_table_minimum_and_smooth_cutoff(p):
    Ei, Ri, ai, Ej, Rj, aj, cr = p
    sig = math.sqrt(4 * Ri * Rj) if cr & 2 else Ri + Rj
    eps = math.sqrt(Ei * Ej) * EUNIT_KCAL
    a = math.sqrt(ai * aj)
    t, corr = table(Provider(oracle.load(), "oracle_"), p)
    assert t.tau == pytest.approx(-a / sig, rel=1e-15)
    assert u_of(t, sig) == pytest.approx(-eps, rel=1e-13)
    assert abs(fr_of(t, sig)) * sig < 1e-12 * eps
    # u(r) is a minimum at sig: above it on both sides
    assert u_of(t, 0.98 * sig) > -eps and u_of(t, 1.02 * sig) > -eps
    C1 = math.sqrt(t.C1q)
    assert 0.5 * C2 < C1 < 0.999 * C2 and t.C2q == C2 * C2
    d = t.C1q - t.C2q
    assert t.A * d * d == pytest.approx(u_of(t, C1), rel=1e-12)
    assert t.A4 * d == pytest.approx(fr_of(t, C1), rel=1e-12)
    assert t.A4 == pytest.approx(-4 * t.A, rel=1e-10)


@pytest.mark.parametrize("p", PAIRS)
def test_tail_correction_against_quadrature(p):
    t, corr = table(Provider(oracle.load(), "oracle_"), p)
    C1 = math.sqrt(t.C1q)
    c0 = -4 * math.pi * integrate.quad(lambda r: t.A * (r * r - t.C2q) ** 2 * r * r, C1, C2,
                                       epsabs=0, epsrel=1e-13)[0]
    c1 = 4 * math.pi * integrate.quad(lambda r: u_of(t, r) * r * r, C1, np.inf,
                                      epsabs=0, epsrel=1e-13)[0]
    assert corr == pytest.approx(c0 + c1, rel=1e-9)


@pytest.mark.parametrize("p", PAIRS)
def test_oracle_and_engine_setss_agree(p):
    provs = providers()
    if len(provs) < 2:
        pytest.skip("engine library not built")
    (t0, c0), (t1, c1) = [table(pr, p) for pr in provs]
    for k in ("tau", "eu", "ef", "ru", "C1q", "C2q", "A", "A4"):
        assert getattr(t1, k) == pytest.approx(getattr(t0, k), rel=1e-14, abs=0), k
    assert c1 == pytest.approx(c0, rel=1e-13)


def test_rejections():
    o = Provider(oracle.load(), "oracle_")
    with pytest.raises(RuntimeError):               # Kirkwood rule
        o.setss_exp6(-50.0, 1.7, 14.0, -50.0, 1.7, 14.0, 0, C2)
    with pytest.raises(RuntimeError):               # positive EvdW
        o.setss_exp6(50.0, 1.7, 14.0, -50.0, 1.7, 14.0, 1, C2)
    t, c = o.setss_exp6(0.0, 1.7, 14.0, -50.0, 1.7, 14.0, 1, C2)
    assert t.C2q == 0 and t.A == 0 and c == 0       # no term: zero table


def neutral_fluid(n=6, seed=11):
    """n^3 exp-6 atoms (two types) on a jittered lattice, no charges"""
    sp = lambda name, t: dict(name=name, site_type=[t], site_charge_e=[0.0], site_xyz=[[0, 0, 0]],
                              site_excl=[[]], bonds=[])
    ff = dict(types=["AR", "KR"], type_mass=[39.948, 83.798], type_EvdW_kcal=[-0.238, -0.325],
              type_RvdW=[1.91, 2.02], type_parm=[13.0, 14.5], form="exp6",
              species=[sp("AR", 0), sp("KR", 1)])
    rng = np.random.default_rng(seed)
    a0 = 4.1
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) * a0
    r = g + rng.uniform(-0.4, 0.4, size=g.shape) + 0.5
    perm = rng.permutation(len(r))
    nA = len(r) // 2
    return ff, [nA, len(r) - nA], n * a0, r[perm]


def test_pair_sum_against_brute_force():
    ff, counts, L, r = neutral_fluid()
    cutoff = 10.0
    bs = build_system(ff, counts, [L] * 3, cutoff, alpha=0.3, kappa=0.3,
                      provider=Provider(oracle.load(), "oracle_"))
    f, out = oracle.forces(bs, r, 1)
    types = bs.info["site_type"]
    nt = 2
    tab = bs.info["sstab"]
    U = vir = 0.0
    fb = np.zeros_like(r)
    pv = np.zeros(6)
    for i in range(len(r) - 1):
        d = r[i] - r[i + 1:]
        d -= L * np.rint(d / L)
        rr = (d * d).sum(1)
        for k in np.nonzero(rr < cutoff * cutoff)[0]:
            t = tab[types[i] * nt + types[i + 1 + k]]
            q = rr[k]
            if q >= t.C2q:
                continue
            if q < t.C1q:
                x = math.sqrt(q)
                u, fr = t.eu * math.exp(t.tau * x) - t.ru / q ** 3, fr_of(t, x)
            else:
                u, fr = t.A * (q - t.C2q) ** 2, t.A4 * (q - t.C2q)
            U += u
            vir += q * fr
            fv = fr * d[k]
            fb[i] += fv
            fb[i + 1 + k] -= fv
            dd = d[k]
            pv += fr * np.array([dd[0] * dd[0], dd[1] * dd[1], dd[2] * dd[2],
                                 dd[1] * dd[2], dd[2] * dd[0], dd[0] * dd[1]])
    assert out.En_pot == pytest.approx(U, rel=1e-12)
    assert out.En_vir == pytest.approx(vir, rel=1e-12)
    assert np.abs(f - fb).max() < 1e-11 * np.abs(fb).max()
    np.testing.assert_allclose(out.Pvir, pv, rtol=1e-11, atol=1e-11 * np.abs(pv).max())
    assert out.nlj > 0


def test_exp6_brine_builds_and_differs_from_lj():
    """the exp-6 water of the workloads: finite forces, the LJ water's
    pair census and Coulomb energy, another site-site energy"""
    wl = workloads.water_c1()
    ffe = workloads.brine_forcefield_exp6(False)
    prov = Provider(oracle.load(), "oracle_")
    bse = build_system(ffe, wl["counts"], [wl["L"]] * 3, wl["cutoff"], provider=prov)
    bsl = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], provider=prov)
    fe, oe = oracle.forces(bse, wl["r"], 1)
    fl, ol = oracle.forces(bsl, wl["r"], 1)
    assert np.isfinite(fe).all() and oe.npair == ol.npair and oe.En_el == ol.En_el
    assert oe.En_pot != ol.En_pot
