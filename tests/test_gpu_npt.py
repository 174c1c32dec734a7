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
