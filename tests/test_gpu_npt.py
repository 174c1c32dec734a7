"""The isotropic MTK and Hoover barostats on the GPU against cook's own
runs (tests/golden/spce216_npt, spce216_hoover: 216 SPC/E at 100 MPa,
tau.T = 0.05, tau.P = 0.2, predictor length 3): from the configuration cook
saved (cfg0) the engine's NPT steps follow the dumped positions and boxes
and end on cook's final configuration, box and barostat variables (cfg1);
the pressure that drives each step equals the oracle's; the NPT state
round-trips through mdx_get_npt / mdx_set_npt with the same bits as an
uninterrupted run; the mixed engine stays within its error bound; bad
parameters are refused."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider
from test_oracle_npt import CASES, load_case, npt_params

pytestmark = pytest.mark.gpu


def start(bs, c0, mode):
    e = Engine(bs, mode)
    e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"])
    e.set_npt(c0["cfgs"][0]["logs"], c0["cfgs"][1]["logs"], c0["cfgs"][1]["lambda"][0], 0, 0)
    return e


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-9), (abi.MDX_MIXED, 2e-5)])
def test_engine_follows_cooks_npt_run(name, mode, tol, tmp_path):
    d, meta, (c0, c1) = load_case(name, tmp_path)
    print("%s mode %d" % (name, mode), flush=True)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    p = npt_params(meta, bs)
    e = start(bs, c0, mode)
    # the oracle beside it, for the pressure of every step
    bso = oracle.fixture_system(meta)
    r_o = c0["cfgs"][0]["rp"].copy()
    r1_o = c0["cfgs"][1]["rp"].copy()
    L_o = np.array(c0["L"], dtype=np.float64)
    L0 = L_o.copy()
    npt_o = np.array([c0["cfgs"][0]["logs"], c0["cfgs"][1]["logs"], c0["cfgs"][1]["lambda"][0],
                      0, 0, 0], dtype=np.float64)
    hist_o = np.zeros((p.trvp_k, 2 + r_o.size))
    kk = 0
    for s in range(meta["nsteps"]):
        r, _ = e.download_state()
        assert np.abs(r - d["r"][s]).max() < tol, s
        assert np.abs(e.box() - d["L"][s]).max() < max(6e-9, tol), s
        so = e.md_steps(p, 1)
        _, kk = oracle.md_step_npt(bso, r_o, r1_o, L_o, L0, npt_o, kk, hist_o, p)
        # the pressure is a sum of virial terms of O(100) program units
        assert abs(so.Pref - npt_o[4]) < 1e3 * tol, (s, so.Pref, npt_o[4])
        assert np.array_equal(np.array(so.L[:]), e.box())
    r, r1 = e.download_state()
    logs, logs1, lam1, k, kk, _ = e.get_npt()
    L = e.box()
    e.close()
    print("  final r %.2e r1 %.2e L %.15g (cook %.15g) logs %.15g (cook %.15g) lambda1 %.15g "
          "(cook %.15g)" % (np.abs(r - c1["cfgs"][0]["rp"]).max(),
                            np.abs(r1 - c1["cfgs"][1]["rp"]).max(), L[0], c1["L"][0], logs,
                            c1["cfgs"][0]["logs"], lam1, c1["cfgs"][1]["lambda"][0]), flush=True)
    assert (k, kk) == (meta["pred"], meta["pred"])
    assert np.abs(r - c1["cfgs"][0]["rp"]).max() < tol
    assert np.abs(L - c1["L"]).max() < (1e-12 if mode == abi.MDX_FP64 else 1e-6) * L[0]
    assert logs == pytest.approx(c1["cfgs"][0]["logs"], rel=100 * tol)
    assert lam1 == pytest.approx(c1["cfgs"][1]["lambda"][0], rel=1000 * tol)
    assert so.logs == logs and np.isfinite(so.En_ext) and so.V > 0


def test_npt_state_round_trip(tmp_path):
    """4 steps, the state read and written back into a fresh context, 4
    more: the same bits as 8 steps in one context"""
    d, meta, (c0, c1) = load_case("spce216_npt", tmp_path)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    p = npt_params(meta, bs)
    ea = start(bs, c0, abi.MDX_FP64)
    oa = ea.md_steps(p, 8)
    eb = start(bs, c0, abi.MDX_FP64)
    eb.md_steps(p, 4)
    st = eb.get_npt()
    r, r1 = eb.download_state()
    L = eb.box()
    eb.close()
    assert st[3:5] == (3, 3) and np.abs(st[5][:, :2]).max() > 0
    ec = Engine(bs, abi.MDX_FP64)
    ec.set_box(L)
    ec.upload_state(r, r1)
    ec.set_npt(*st)
    assert np.array_equal(ec.get_npt()[5], st[5])
    oc = ec.md_steps(p, 4)
    (ra, r1a), (rc, r1c) = ea.download_state(), ec.download_state()
    assert np.array_equal(ra, rc) and np.array_equal(r1a, r1c)
    assert np.array_equal(ea.box(), ec.box())
    assert ea.get_npt()[:3] == ec.get_npt()[:3]
    assert oa.En_ext == oc.En_ext and oa.Pref == oc.Pref
    ea.close()
    ec.close()


def test_npt_refuses_bad_parameters(tmp_path):
    d, meta, (c0, c1) = load_case("spce216_npt", tmp_path)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    p = npt_params(meta, bs)
    e = start(bs, c0, abi.MDX_FP64)
    for field, value in (("tau_P", 0.0), ("virial", 2), ("trvp_k", 10), ("P", float("nan"))):
        bad = abi.Step.from_buffer_copy(p)
        setattr(bad, field, value)
        with pytest.raises(RuntimeError):
            e.md_steps(bad, 1)
    with pytest.raises(RuntimeError):
        e.set_npt(0, 0, 0, 3, 4)
    # nothing moved
    r, _ = e.download_state()
    assert np.array_equal(r, c0["cfgs"][0]["rp"]) and np.array_equal(e.box(), c0["L"])
    e.close()


def test_extended_energy_fluctuation_is_second_order_in_h(tmp_path):
    """the conserved quantity of the MTK step, En.tot = En.U + En.ext
    (sim/constrd.c:1060-1068, sim/thermo.c:69-84, with the cutoff
    correction Ecorr/V of the current box), over 2 ps from cook's start
    at h = 2 fs and 1 fs: its fluctuation is that of the leap-frog energy
    estimate, second order in h (a wrong barostat force would leave an
    h-independent error), and it does not drift by more than it fluctuates;
    Nose-Hoover at 2 fs from the same start fluctuates as much"""
    d, meta, (c0, c1) = load_case("spce216_npt", tmp_path)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    p = npt_params(meta, bs)

    def series(h, th):
        e = Engine(bs, abi.MDX_FP64)
        e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"] * (h / meta["h"]))
        q = abi.Step.from_buffer_copy(p)
        q.h, q.thermostat = h, th
        if th == 2:
            e.set_nose(0, 0, 0, 0)
        else:
            e.set_npt(0, 0, 0, 0, 0)
        n = int(round(2.0 / h))
        blk = n // 200
        H = []
        for _ in range(200):
            so = e.md_steps(q, blk)
            V = so.V if th != 2 else meta["L"] ** 3
            H.append(so.En_kin + so.frc.En_pot + so.frc.En_el + p.Ecorr / V + so.En_ext)
        L = e.box()[0]
        e.close()
        H = np.array(H)
        t = np.arange(len(H))
        drift = np.polyval(np.polyfit(t, H, 1), [0, len(H) - 1])
        return H.std(), abs(drift[1] - drift[0]), L

    s2, d2, L2 = series(0.002, 22)
    s1, d1, L1 = series(0.001, 22)
    sn, dn, _ = series(0.002, 2)
    print("En.tot std %.1f K at 2 fs, %.1f K at 1 fs (ratio %.2f), drift over 2 ps %.1f / %.1f K; "
          "Nose-Hoover std %.1f K; L %.4f / %.4f" % (s2, s1, s2 / s1, d2, d1, sn, L2, L1), flush=True)
    assert 3.0 < s2 / s1 < 5.5
    assert d2 < 2 * s2 and d1 < 2 * s1
    assert 0.5 < s2 / sn < 2
    assert abs(L2 / L1 - 1) < 0.02
