"""The charge adjoint (mdx_adjoint_charges) against exact differences of the
force call: every force and energy term is bilinear in the charges, so for
a direction dq of the site charges the central difference
[G(q + h dq) - G(q - h dq)] / 2h of G = sum_i w_i . F_i (and of En_el) is
the exact directional derivative up to rounding.  Checked on cook's
fixtures with tin-foil, net-charge background, Yeh-Berkowitz slab and
both corrections, in directions that keep and that change the net charge,
against the engine's own FP64 force call (to 1e-10) and against the oracle's
(to 1e-8); a mixed-mode context gives the same adjoint (FP64 kernels)."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu

CASES = ["brine506", "spce1000_md", "brine505_charged", "brine506_slab", "brine505_slabbg"]


def directions(bs, rng):
    """random site directions (the net charge changes) and a per-sign
    direction (negative sites +1, positive -1/2: neutral for water), on the
    charged sites"""
    q0 = bs.info["charge"]
    ch = (q0 != 0).astype(float)
    out = [rng.normal(size=q0.size) * ch * np.abs(q0).max() for _ in range(2)]
    out.append(np.where(q0 < 0, 1.0, -0.5) * ch * np.abs(q0).max())
    return out


def central(fun, q0, dq, h):
    gp, ep = fun(q0 + h * dq)
    gm, em = fun(q0 - h * dq)
    return (gp - gm) / (2 * h), (ep - em) / (2 * h)


@pytest.mark.parametrize("name", CASES)
def test_charge_adjoint_vs_exact_differences(name):
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    r = d["r"]
    ns = bs.struct.ns
    rng = np.random.default_rng(7)
    w = rng.normal(size=(ns, 3))
    eng = Engine(bs, abi.MDX_FP64)
    gF, gE = eng.adjoint_charges(r, w)
    q0 = bs.info["charge"].copy()
    assert np.all(gF[q0 == 0] == 0) and np.all(gE[q0 == 0] == 0)

    def engine_G(q):
        eng.set_params(charge=q)
        f, out = eng.forces(r, measure=1)
        return float((w * f).sum()), out.En_el

    bso = oracle.fixture_system(meta)
    qo = bso.info["charge"]          # the array the oracle's system points into

    def oracle_G(q):
        qo[:] = q
        f, out = oracle.forces(bso, r)
        return float((w * f).sum()), out.En_el

    h = 0.01
    for k, dq in enumerate(directions(bs, rng)):
        pF, pE = float(gF @ dq), float(gE @ dq)
        fF, fE = central(engine_G, q0, dq, h)
        oF, oE = central(oracle_G, q0, dq, h)
        eF, eE = abs(pF - fF) / abs(fF), abs(pE - fE) / abs(fE)
        print("%s dir %d: dG %.12e adjoint vs engine diff %.1e, dE %.12e %.1e; vs oracle %.1e %.1e"
              % (name, k, fF, eF, fE, eE, abs(pF - oF) / abs(oF), abs(pE - oE) / abs(oE)), flush=True)
        assert eF < 1e-10 and eE < 1e-10
        assert abs(pF - oF) < 1e-8 * abs(oF) and abs(pE - oE) < 1e-8 * abs(oE)
    qo[:] = q0
    eng.set_params(charge=q0)
    # the same adjoint from a context of the mixed mode
    em = Engine(bs, abi.MDX_MIXED)
    gF32, gE32 = em.adjoint_charges(r, w)
    em.close()
    eng.close()
    np.testing.assert_allclose(gF32, gF, rtol=1e-12, atol=1e-12 * np.abs(gF).max())
    np.testing.assert_allclose(gE32, gE, rtol=1e-12, atol=1e-12 * np.abs(gE).max())


def test_charge_adjoint_refuses_drudes():
    d, meta = oracle.load_fixture("polar306")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, abi.MDX_FP64)
    ns = bs.struct.ns
    r = np.concatenate([d["r"][:ns], np.zeros((ns, 3))])     # Drudes on their centres
    with pytest.raises(RuntimeError, match="Drude"):
        eng.adjoint_charges(r, np.zeros((ns, 3)))
    eng.close()
