"""GPU parity of the Drude (POLAR) path through the C ABI: the four-term
pair kernel and Drude charges in Ewald (mdx_forces), one selffield()
iteration (mdx_selffield) and the device-resident ASPC + SCF + SHAKE step
(mdx_md_steps), against the reference cook's dumps of a polarizable
SPC/E + NaCl box (tests/golden/polar306*) and the oracle."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import PUNIT, engine_provider

pytestmark = pytest.mark.gpu

# relative RMS force error bound vs the FP64 reference
TOL = {abi.MDX_FP64: 1e-10, abi.MDX_MIXED: 1e-5}
MODES = [abi.MDX_FP64, abi.MDX_MIXED]


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def pol():
    d, meta = oracle.load_fixture("polar306")
    return d, meta, oracle.fixture_system(meta, provider_=engine_provider())


@pytest.mark.parametrize("mode", MODES)
def test_polar_forces_vs_golden(pol, mode):
    d, meta, bs = pol
    print("polar306 forces mode", mode, flush=True)
    eng = Engine(bs, mode)
    for k in range(d["rpol"].shape[0]):
        r = np.concatenate([d["r"], d["rpol"][k]])
        fr = np.concatenate([d["f"][k], d["fpol"][k]])
        # measure=0 is the FP32 pair kernel in the mixed mode; measured calls
        # of a Drude system take the FP64 pair kernel (k-space stays mixed)
        for measure in (0, 1):
            f, out = eng.forces(r, measure=measure)
            tol = TOL[mode]
            e = rel_rms(f, fr)
            print("dump", k, "measure", measure, "rel force err", e, flush=True)
            assert e < tol, (k, e)
            # the Drude rows alone
            assert rel_rms(f[bs.struct.ns:], d["fpol"][k]) < 10 * tol
    eng.close()


@pytest.mark.parametrize("mode", MODES)
def test_polar_scf_energy_and_tensor(pol, mode):
    """scf_iterations forces()+selffield() from dump 0; the Drude offsets
    after each iteration are the next dump, and the energy and pressure
    tensor after the last are cook's printed ones"""
    d, meta, bs = pol
    ns = bs.struct.ns
    eng = Engine(bs, mode)
    r = np.concatenate([d["r"], d["rpol"][0]])
    for k in range(meta["scf_iterations"]):
        f, out = eng.forces(r, measure=1)
        es, me = eng.selffield(r, f, out, meta["scf_omega"])
        ptol = 1e-10 if mode == abi.MDX_FP64 else 1e-5 * np.abs(d["rpol"][k + 1]).max() * 10
        err = np.abs(r[ns:] - d["rpol"][k + 1]).max()
        print("iteration", k, "rpol err", err, "Eself", es, flush=True)
        assert err < max(ptol, 1e-10), (k, err)
    Epot = out.En_el + out.En_pot + es
    etol = 1e-10 if mode == abi.MDX_FP64 else 1e-5
    assert abs(Epot - meta["Epot"]) < etol * abs(meta["Epot"])
    V = meta["L"] ** 3
    ref = meta["Ptens_printed_Pa"]
    P = [out.Pvir[i] / (3 * V) * PUNIT for i in range(3)] + [out.Pvir[i] * PUNIT for i in range(3, 6)]
    scale = max(abs(x) for x in ref[:3])
    ttol = 1e-9 if mode == abi.MDX_FP64 else 1e-5
    for a, b in zip(P[:3], ref[:3]):
        assert abs(a - b) < ttol * scale
    for a, b in zip(P[3:], ref[3:]):
        assert abs(a - b) < ttol * scale * 3 * V
    eng.close()


@pytest.mark.parametrize("mode", MODES)
def test_polar_selffield_matches_oracle(pol, mode):
    d, meta, bs = pol
    bso = oracle.fixture_system(meta)
    eng = Engine(bs, mode)
    r = np.concatenate([d["r"], d["rpol"][1]])
    f, out = eng.forces(r, measure=1)
    ro, fo = r.copy(), f.copy()
    outo = abi.Out()
    outo.En_el, outo.En_pot = out.En_el, out.En_pot
    outo.Pvir = out.Pvir
    es_o, me_o = oracle.selffield(bso, ro, fo, outo, 0.7)
    es, me = eng.selffield(r, f, out, 0.7)
    # same inputs: the selffield kernel is FP64 in both modes
    assert np.abs(r - ro).max() < 1e-12
    assert rel_rms(f, fo) < 1e-13
    assert abs(es - es_o) < 1e-11 * abs(es_o)
    assert abs(me - me_o) < 1e-9 * abs(me_o)
    assert abs(out.En_el - outo.En_el) < 1e-11 * abs(outo.En_el)
    eng.close()


def _md_params(meta, **kw):
    p = abi.Step(h=meta["h"], eps=1e-12, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=1.0,
                 measure_every=1, aspc_k=oracle.ASPC_K, scf_iter=1, scf_omega=0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("mode", MODES)
def test_polar_md_step_matches_reference(mode):
    """one device-resident step from the state before dump 7 of cook's -p420
    run: ASPC prediction with the full k = 4 history, one SCF iteration,
    SHAKE; the sites land on dump 8"""
    d, meta = oracle.load_fixture("polar306_aspc")
    bso = oracle.fixture_system(meta)
    post = oracle.aspc_post_scf(bso, d)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    ns = bs.struct.ns
    r, r1, hist = oracle.aspc_state(d, post, 7)
    eng = Engine(bs, mode)
    eng.upload_state(r, r1)
    eng.set_aspc(oracle.ASPC_K, oracle.ASPC_K, hist)
    out = eng.md_steps(_md_params(meta), 1)
    rg, r1g = eng.download_state()
    # oracle step from the same state
    ro, r1o, ho = r.copy(), r1.copy(), hist.copy()
    outo, kk = oracle.md_step_polar(bso, ro, r1o, ho, oracle.ASPC_K, _md_params(meta))
    k, kkg, hg = eng.get_aspc()
    assert (k, kkg) == (oracle.ASPC_K, oracle.ASPC_K)
    assert np.array_equal(hg, ho)            # history shift is a copy
    assert np.array_equal(r1g[ns:], r1o[ns:])
    ptol = 1e-10 if mode == abi.MDX_FP64 else 1e-7
    e_ref = np.abs(rg[:ns] - d["r"][8]).max()
    e_pol = np.abs(rg[ns:] - post[7]).max()
    print("mode", mode, "site err", e_ref, "Drude err", e_pol, flush=True)
    assert e_ref < ptol
    assert e_pol < (1e-10 if mode == abi.MDX_FP64 else 1e-4 * np.abs(post[7]).max())
    etol = 1e-10 if mode == abi.MDX_FP64 else 1e-5
    assert abs(out.En_self - outo.En_self) < etol * abs(outo.En_self)
    Eg = out.frc.En_el + out.frc.En_pot
    Eo = outo.frc.En_el + outo.frc.En_pot
    assert abs(Eg - Eo) < etol * abs(Eo)
    ktol = 1e-9 if mode == abi.MDX_FP64 else 1e-6
    assert abs(out.En_kin - outo.En_kin) < ktol * abs(outo.En_kin)
    eng.close()


def test_polar_md_ramp_matches_oracle():
    """six steps from a fresh predictor (kk = 0, zero history, like the
    start of a cook run): the predictor length and omega ramp of
    sim/pred2.c:67-98 in the engine and the oracle agree step by step"""
    d, meta = oracle.load_fixture("polar306_aspc")
    bso = oracle.fixture_system(meta)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    ns = bs.struct.ns
    r = np.concatenate([d["r"][1], d["rpol"][1]])
    r1 = np.concatenate([d["r"][1] - d["r"][0], d["rpol"][1]])
    eng = Engine(bs, abi.MDX_FP64)
    eng.upload_state(r, r1)
    eng.set_aspc(oracle.ASPC_K, 0, None)
    ro, r1o = r.copy(), r1.copy()
    ho = np.zeros((oracle.ASPC_K, ns, 3))
    kk = 0
    for s in range(6):
        out = eng.md_steps(_md_params(meta), 1)
        outo, kk = oracle.md_step_polar(bso, ro, r1o, ho, kk, _md_params(meta))
        rg, r1g = eng.download_state()
        e = np.abs(rg - ro).max()
        print("step", s, "kk", kk, "err", e, flush=True)
        assert e < 1e-9, (s, e)
        assert abs(out.En_self - outo.En_self) < 1e-9 * abs(outo.En_self)
    assert eng.get_aspc()[1] == kk == oracle.ASPC_K
    eng.close()


def test_polar_md_mixed_matches_oracle():
    """six steps in one call, mixed mode: the unmeasured steps run the FP32
    four-term kernel, the last (measured) one the FP64 kernel"""
    d, meta = oracle.load_fixture("polar306_aspc")
    bso = oracle.fixture_system(meta)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    ns = bs.struct.ns
    r = np.concatenate([d["r"][1], d["rpol"][1]])
    r1 = np.concatenate([d["r"][1] - d["r"][0], d["rpol"][1]])
    eng = Engine(bs, abi.MDX_MIXED)
    eng.upload_state(r, r1)
    eng.set_aspc(oracle.ASPC_K, 0, None)
    out = eng.md_steps(_md_params(meta, measure_every=0), 6)
    rg, r1g = eng.download_state()
    ro, r1o = r.copy(), r1.copy()
    ho = np.zeros((oracle.ASPC_K, ns, 3))
    kk = 0
    for s in range(6):
        outo, kk = oracle.md_step_polar(bso, ro, r1o, ho, kk, _md_params(meta))
    e_site = np.abs(rg[:ns] - ro[:ns]).max()
    e_pol = np.abs(rg[ns:] - ro[ns:]).max()
    print("site err", e_site, "Drude err", e_pol, flush=True)
    # six steps of h^2 dF/m with |dF| <= 1e-5 |F|
    assert e_site < 1e-6
    assert e_pol < 1e-4 * np.abs(ro[ns:]).max()
    Eg = out.frc.En_el + out.frc.En_pot + out.En_self
    Eo = outo.frc.En_el + outo.frc.En_pot + outo.En_self
    assert abs(Eg - Eo) < 1e-5 * abs(Eo)
    eng.close()


def test_polar_bad_drude_is_an_error(pol):
    """a Drude offset beyond el.minqq (1 A) is rejected (sim/scf.c), from
    the host-side check of the kernel's flag"""
    d, meta, bs = pol
    eng = Engine(bs, abi.MDX_FP64)
    r = np.concatenate([d["r"], d["rpol"][0]])
    i = int(np.nonzero(bs.info["chargepol"])[0][0])
    r[bs.struct.ns + i] = [1.5, 0.0, 0.0]
    f, out = eng.forces(r, measure=1)
    with pytest.raises(RuntimeError):
        eng.selffield(r, f, out, 0.7)
    eng.close()
