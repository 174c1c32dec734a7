"""The Busing form (sim/busing, POW = 0),
u = f rho_ij exp((R_ij - r)/rho_ij) - C_ij/r^6, f = 0.05 e^2/A^2, in the
exp-6 table; pinned by what the forms fix (cook's Busing build needs the
generated simopt.h, like exp-6 and Buckingham):

- f is 8355.047344920979 K/A, the value sim/busing/sitesite.h:11 states;
- u_rep(R_ij) = f rho_ij = EvdW_i + EvdW_j, and u_rep'/u_rep = -1/rho_ij;
- Busing sites made from Buckingham sites by the conversion of
  sim/busing/sitesite.h:24-26 (rho = 1/2B, R = rho ln(A/2 f rho),
  C_i = sqrt(C)) give the Buckingham table of additive 1/B and the Busing
  rule for A (comb_rule 15) for like and unlike pairs, with the same
  smooth cutoff and tail correction;
- a site of EvdW < 0, or a pair of EvdW = 0, is switched off whole (no
  C/r^6 either); comb_rule other than 0 is refused;
- the oracle's and the engine's set-up helpers agree.
"""
import math

import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.system import Provider

C2 = 9.0
F_BUSING = 0.05 * 408.779826922048 ** 2
BUCK = [(3.0e7, 3.6, 1.2e6), (1.5e8, 3.1, 3.3e6), (8.0e6, 4.2, 2.0e5)]  # A [K], B [1/A], C [K A^6]


def as_busing(A, B, C):
    """sim/busing/sitesite.h:24-26: EvdW = f rho [K], R [A], C_i [sqrt(K) A^3]"""
    rho = 0.5 / B
    return F_BUSING * rho, rho * math.log(A / (2 * F_BUSING * rho)), math.sqrt(C)


def provs():
    out = [Provider(oracle.load(), "oracle_")]
    try:
        out.append(Provider(abi.load_engine(), "mdx_"))
    except (OSError, RuntimeError):
        pass
    return out


def test_busing_constant():
    assert F_BUSING == pytest.approx(8355.047344920979, rel=1e-15)


@pytest.mark.parametrize("p", range(2))
def test_repulsion_at_R_is_f_rho(p):
    prov = provs()[p] if p < len(provs()) else pytest.skip("engine library not built")
    Ei, Ri, Ci = 520.0, 1.7, 900.0
    Ej, Rj, Cj = 260.0, 1.1, 1500.0
    t, _ = prov.setss_busing(Ei, Ri, Ci, Ej, Rj, Cj, 0, C2)
    rho = (Ei + Ej) / F_BUSING
    assert t.eu * math.exp(t.tau * (Ri + Rj)) == pytest.approx(Ei + Ej, rel=1e-14)
    assert t.tau == pytest.approx(-1 / rho, rel=1e-15)
    assert t.ef == pytest.approx(t.eu / rho, rel=1e-15)
    assert t.ru == pytest.approx(Ci * Cj, rel=1e-15)


@pytest.mark.parametrize("i,j", [(0, 0), (1, 1), (0, 1), (1, 2), (0, 2)])
def test_buckingham_sites_converted(i, j):
    for prov in provs():
        bk, cb = prov.setss_buck(*BUCK[i], *BUCK[j], 15, C2)
        bu, cu = prov.setss_busing(*as_busing(*BUCK[i]), *as_busing(*BUCK[j]), 0, C2)
        for k in ("tau", "eu", "ef", "ru", "C1q", "C2q", "A", "A4"):
            assert getattr(bu, k) == pytest.approx(getattr(bk, k), rel=1e-12), k
        assert cu == pytest.approx(cb, rel=1e-11)
        assert bu.A != 0 and bu.C1q < bu.C2q


def test_switched_off_pairs_and_rules():
    o = Provider(oracle.load(), "oracle_")
    for prov in provs():
        for Ei, Ej in ((-1.0, 500.0), (500.0, -1.0), (0.0, 0.0)):
            t, c = prov.setss_busing(Ei, 1.5, 900.0, Ej, 1.5, 900.0, 0, C2)
            assert (t.tau, t.eu, t.ef, t.ru, t.A, t.A4, t.C1q, t.C2q, c) == (0,) * 9
        # one bare site (EvdW = 0) still combines with another
        t, _ = prov.setss_busing(0.0, 0.0, 30.0, 500.0, 1.5, 30.0, 0, C2)
        assert t.eu > 0 and t.ru == 900.0
        with pytest.raises(RuntimeError):
            prov.setss_busing(500.0, 1.5, 900.0, 500.0, 1.5, 900.0, 1, C2)
    t, c = o.setss_busing(*as_busing(*BUCK[0]), *as_busing(*BUCK[1]), 0, C2)
    assert c < 0                      # the tail is dominated by -C/r^6


def test_oracle_and_engine_agree():
    pv = provs()
    if len(pv) < 2:
        pytest.skip("engine library not built")
    for i in range(3):
        for j in range(i + 1):
            a = [pv[k].setss_busing(*as_busing(*BUCK[i]), *as_busing(*BUCK[j]), 0, C2)
                 for k in range(2)]
            for k in ("tau", "eu", "ef", "ru", "C1q", "C2q", "A", "A4"):
                assert getattr(a[0][0], k) == pytest.approx(getattr(a[1][0], k), rel=1e-14)
            assert a[0][1] == pytest.approx(a[1][1], rel=1e-13)


def test_brine_forces_equal_the_buckingham_system():
    """the oracle's force call on cook's brine506 positions: the converted
    Busing sites give the forces and energy of the Buckingham system
    (comb_rule 15), pair for pair (same census)"""
    import numpy as np
    from macsimus_amd import workloads
    from macsimus_amd.system import EUNIT_KCAL
    d, meta = oracle.load_fixture("brine506")
    base = meta["forcefield"]
    bk, bu = [], []
    for t, Ev, Rv in zip(base["types"], base["type_EvdW_kcal"], base["type_RvdW"]):
        a, sig, eps = workloads.EXP6_A[t], 2 * Rv, -Ev * EUNIT_KCAL
        site = (6 * eps * math.exp(a) / (a - 6), a / sig, eps * a * sig ** 6 / (a - 6))
        bk.append(site)
        bu.append(as_busing(*site) if site[0] > 0 else (-1.0, 0.0, math.sqrt(site[2])))
    ffk = dict(base, form="buck", comb_rule=15,
               **{k: [s[n] / EUNIT_KCAL ** p for s in bk] for k, n, p in
                  (("type_EvdW_kcal", 0, 1), ("type_RvdW", 1, 0), ("type_parm", 2, 1))})
    ffb = dict(base, form="busing", comb_rule=0,
               **{k: [s[n] / EUNIT_KCAL ** p for s in bu] for k, n, p in
                  (("type_EvdW_kcal", 0, 1), ("type_RvdW", 1, 0), ("type_parm", 2, 0.5))})
    fk, ok = oracle.forces(oracle.fixture_system(dict(meta, forcefield=ffk)), d["r"], 1)
    fb, ob = oracle.forces(oracle.fixture_system(dict(meta, forcefield=ffb)), d["r"], 1)
    assert ob.nlj == ok.nlj > 0 and ob.npair == ok.npair
    assert np.sqrt(((fb - fk) ** 2).sum() / (fk ** 2).sum()) < 1e-13
    assert ob.En_pot == pytest.approx(ok.En_pot, rel=1e-12)
    assert ob.En_vir == pytest.approx(ok.En_vir, rel=1e-12)
