"""Drude (POLAR) restatement in the oracle against the reference cook's
dumps of a polarizable SPC/E + NaCl box (tests/golden/polar306.*): the
four-term pair kernel, Drude charges in Ewald, and selffield()."""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def pol():
    d, meta = oracle.load_fixture("polar306")
    return d, meta, oracle.fixture_system(meta)


def test_setup(pol):
    d, meta, bs = pol
    assert bs.info["nk"] == meta["Nk"] and bs.info["ncell"] == meta["ncell"]
    assert abs(bs.info["alpha"] - meta["alpha_printed"]) < 1e-6


def test_forces_on_centres_and_drudes(pol):
    d, meta, bs = pol
    for k in range(d["rpol"].shape[0]):
        r = np.concatenate([d["r"], d["rpol"][k]])
        f, out = oracle.forces(bs, r)
        fr = np.concatenate([d["f"][k], d["fpol"][k]])
        rel = np.sqrt(((f - fr) ** 2).sum() / (fr ** 2).sum())
        assert rel < 1e-10, (k, rel)


def test_selffield_moves_drudes_like_reference(pol):
    """dumps 0..3 are consecutive SCF iterations with omega=0.7"""
    d, meta, bs = pol
    ns = bs.struct.ns
    for k in range(3):
        r = np.concatenate([d["r"], d["rpol"][k]])
        f, out = oracle.forces(bs, r)
        es, me = oracle.selffield(bs, r, f, out, 0.7)
        assert np.abs(r[ns:] - d["rpol"][k + 1]).max() < 1e-10
        assert es > 0


ASPC_K = oracle.ASPC_K


@pytest.fixture(scope="module")
def aspc():
    d, meta = oracle.load_fixture("polar306_aspc")
    bs = oracle.fixture_system(meta)
    return d, meta, bs, oracle.aspc_post_scf(bs, d)


def _state(d, bs, post, k):
    return oracle.aspc_state(d, post, k)


def test_aspc_prediction_matches_reference(aspc):
    d, meta, bs, post = aspc
    ns = bs.struct.ns
    for k in (7, 8):
        r, r1, hist = _state(d, bs, post, k)
        p = oracle.abi.Step(h=meta["h"], eps=1e-12, aspc_k=ASPC_K, scf_iter=-1)
        out, kk = oracle.md_step_polar(bs, r, r1, hist, ASPC_K, p)
        assert kk == ASPC_K
        assert np.abs(r[ns:] - d["rpol"][k]).max() < 1e-10, k
        # history shifted by one: aa[1] <- aa[0], aa[2] <- aa[1], ...
        assert np.array_equal(r1[ns:], post[k - 1])
        assert np.array_equal(hist[0], post[k - 2])


def test_aspc_predictor_coefficients_are_exact_for_polynomials(aspc):
    """B(kk) reproduces a polynomial of degree <= 1 exactly and the k=0
    predictor is 2 a0 - a1 (sim/pred2.c:48); a random-walk check would
    not see a wrong B, a linear ramp does"""
    d, meta, bs, post = aspc
    ns = bs.struct.ns
    rng = np.random.default_rng(5)
    base = rng.normal(size=(ns, 3))
    slope = rng.normal(size=(ns, 3))
    for kk in range(0, ASPC_K + 1):
        r = np.concatenate([d["r"][0], base])
        r1 = np.concatenate([np.zeros((ns, 3)), base - slope])
        hist = np.stack([base - (m + 2) * slope for m in range(ASPC_K)])
        p = oracle.abi.Step(h=meta["h"], eps=1e-12, aspc_k=ASPC_K, scf_iter=-1)
        out, kk1 = oracle.md_step_polar(bs, r, r1, hist, kk, p)
        assert kk1 == min(kk + 1, ASPC_K)
        assert np.abs(r[ns:] - (base + slope)).max() < 1e-12 * (1 + np.abs(base).max()) * 50


def test_aspc_md_step_matches_reference_trajectory(aspc):
    d, meta, bs, post = aspc
    ns = bs.struct.ns
    r, r1, hist = _state(d, bs, post, 7)
    p = oracle.abi.Step(h=meta["h"], eps=1e-12, T=298.0, tau=0.1, nf=1.0,
                        aspc_k=ASPC_K, scf_iter=1, measure_every=1)
    out, kk = oracle.md_step_polar(bs, r, r1, hist, ASPC_K, p)
    assert np.abs(r[:ns] - d["r"][8]).max() < 1e-10
    assert np.abs(r[ns:] - post[7]).max() < 1e-10
    assert out.En_self > 0 and out.frc.npair > 0


def test_energy_and_pressure_tensor_after_scf(pol):
    """cook prints Epot = En.el + En.pot + En.self and the pressure tensor
    after the last of scf_iterations selffield() calls (sim/scf.c:12-240,
    cook/forces.c); dump k is the state before iteration k+1"""
    from macsimus_amd.system import PUNIT
    d, meta, bs = pol
    k = meta["scf_iterations"] - 1
    r = np.concatenate([d["r"], d["rpol"][k]])
    f, out = oracle.forces(bs, r)
    es, me = oracle.selffield(bs, r, f, out, meta["scf_omega"])
    Epot = out.En_el + out.En_pot + es
    assert abs(Epot - meta["Epot"]) < 1e-10 * abs(meta["Epot"])
    V = meta["L"] ** 3
    ref = meta["Ptens_printed_Pa"]
    P = [out.Pvir[i] / (3 * V) * PUNIT for i in range(3)] + [out.Pvir[i] * PUNIT for i in range(3, 6)]
    scale = max(abs(x) for x in ref[:3])
    for a, b in zip(P[:3], ref[:3]):
        assert abs(a - b) < 1e-9 * scale
    for a, b in zip(P[3:], ref[3:]):
        assert abs(a - b) < 1e-9 * scale * 3 * V
