"""Full-size checks (194,304-site SPC/E + NaCl, 12 A cutoff, Ewald at
0.05 K/A): the mixed engine against the FP64 GPU path (itself pinned to
the reference at 1e-10 on the golden systems), and size-independent
properties: momentum conservation (sum of forces), invariance of forces,
energy and virial under a rigid translation of all sites."""
import numpy as np
import pytest

from macsimus_amd import abi, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import build_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    wl = workloads.brine_c2()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    res = {}
    for mode in (abi.MDX_FP64, abi.MDX_MIXED):
        eng = Engine(bs, mode)
        f, out = eng.forces(wl["r"], measure=1)
        shift = np.array([3.7, -11.2, 0.45])
        f2, out2 = eng.forces(wl["r"] + shift, measure=1)
        res[mode] = (f, out, f2, out2)
        eng.close()
    return wl, bs, res


def rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_mixed_force_error_vs_fp64(c2):
    wl, bs, res = c2
    f64, o64 = res[abi.MDX_FP64][:2]
    f32, o32 = res[abi.MDX_MIXED][:2]
    e = rel(f32, f64)
    print("C2 mixed vs FP64: rel RMS force err %.3e" % e, flush=True)
    assert e < 1e-5
    assert o32.npair == o64.npair
    E64 = o64.En_pot + o64.En_el
    E32 = o32.En_pot + o32.En_el
    print("C2 Epot fp64 %.10e mixed %.10e rel %.2e" % (E64, E32, abs(E32 - E64) / abs(E64)), flush=True)
    assert abs(E32 - E64) < 1e-6 * abs(E64)
    for i in range(6):
        scale = max(abs(o64.Pvir[k]) for k in range(3))
        assert abs(o32.Pvir[i] - o64.Pvir[i]) < 1e-5 * scale


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_total_force_vanishes(c2, mode):
    wl, bs, res = c2
    f = res[mode][0]
    fmax = np.abs(f).max()
    tol = 1e-9 if mode == abi.MDX_FP64 else 2e-5
    assert np.abs(f.sum(0)).max() < tol * fmax * np.sqrt(len(f))


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_translation_invariance(c2, mode):
    wl, bs, res = c2
    f, out, f2, out2 = res[mode]
    tol = 1e-9 if mode == abi.MDX_FP64 else 1e-5
    assert rel(f2, f) < tol
    assert out2.npair == out.npair
    E, E2 = out.En_pot + out.En_el, out2.En_pot + out2.En_el
    assert abs(E2 - E) < tol * abs(E)


@pytest.fixture(scope="module")
def c3():
    """the Drude-polarizable brine (97,200 sites + 32,600 Drudes, q_D = -8 e)
    after 3 SCF iterations from Drudes on their centres, in FP64: a state
    with nonzero Drude offsets for both precisions to evaluate"""
    wl = workloads.polar_c3()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    eng = Engine(bs, abi.MDX_FP64)
    r = np.concatenate([wl["r"], np.zeros_like(wl["r"])])
    for _ in range(3):
        f, out = eng.forces(r, measure=1)
        eng.selffield(r, f, out, 0.7)
    eng.close()
    return wl, bs, r


def test_polar_c3_mixed_vs_fp64(c3):
    wl, bs, r = c3
    ns = bs.struct.ns
    assert np.abs(r[ns:]).max() > 1e-3          # the Drudes moved
    res = {}
    for mode in (abi.MDX_FP64, abi.MDX_MIXED):
        eng = Engine(bs, mode)
        if mode == abi.MDX_MIXED:
            # Drudes: pass 1 on the FP32 matrix cores by default
            assert eng.info()["kpass1_tiles"] == 2
        res[mode] = [eng.forces(r, measure=m) for m in (0, 1)]
        eng.close()
    # measure=0: FP32 four-term pair kernel in the mixed mode
    for m in (0, 1):
        f64, o64 = res[abi.MDX_FP64][m]
        f32, o32 = res[abi.MDX_MIXED][m]
        e = rel(f32, f64)
        ed = rel(f32[ns:], f64[ns:])
        print("C3 measure=%d mixed vs FP64: rel RMS force err %.3e (Drude rows %.3e)" % (m, e, ed), flush=True)
        assert e < 1e-5 and ed < 1e-5
        assert o32.npair == o64.npair
    # measured call: energy and tensor
    o64, o32 = res[abi.MDX_FP64][1][1], res[abi.MDX_MIXED][1][1]
    E64, E32 = o64.En_pot + o64.En_el, o32.En_pot + o32.En_el
    print("C3 Epot fp64 %.10e mixed %.10e rel %.2e" % (E64, E32, abs(E32 - E64) / abs(E64)), flush=True)
    assert abs(E32 - E64) < 1e-5 * abs(E64)
    scale = max(abs(o64.Pvir[k]) for k in range(3))
    for i in range(6):
        assert abs(o32.Pvir[i] - o64.Pvir[i]) < 1e-5 * scale


def test_polar_c3_md_steps(c3):
    """a few device-resident ASPC + SCF + SHAKE steps in both precisions
    stay together"""
    wl, bs, r = c3
    out = {}
    for mode in (abi.MDX_FP64, abi.MDX_MIXED):
        eng = Engine(bs, mode)
        eng.upload_state(r, np.zeros_like(r))
        p = abi.Step(h=0.001, eps=1e-7, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=1.0,
                     measure_every=0, aspc_k=4, scf_iter=1, scf_omega=0.0)
        so = eng.md_steps(p, 5)
        rr, r1 = eng.download_state()
        out[mode] = (so, rr)
        eng.close()
    (s64, r64), (s32, r32) = out[abi.MDX_FP64], out[abi.MDX_MIXED]
    ns = bs.struct.ns
    d = np.abs(r32[:ns] - r64[:ns]).max()
    print("C3 5 steps: max site deviation %.3e A, Eself %.6e / %.6e" % (d, s64.En_self, s32.En_self), flush=True)
    assert d < 1e-6
    assert abs(s32.En_self - s64.En_self) < 1e-5 * abs(s64.En_self)
    assert s64.shake_iter_max > 0


def _vs_oracle(bs, bso, r, label, polar=False):
    import oracle
    fo, oo = oracle.forces(bso, r, 1, nthreads=16)
    for mode, tol, etol in ((abi.MDX_FP64, 1e-10, 1e-10), (abi.MDX_MIXED, 1e-5, 2e-6)):
        e = Engine(bs, mode)
        f, out = e.forces(r, 1)
        e.close()
        err = rel(f, fo)
        Etot = oo.En_pot + oo.En_el
        dE = out.En_pot + out.En_el - Etot
        print("%s mode %d vs oracle: rel RMS force err %.3e, Epot rel %.2e, pairs %d" % (
            label, mode, err, dE / Etot, out.npair), flush=True)
        assert err < tol
        assert out.npair == oo.npair
        if mode == abi.MDX_FP64 and not polar:
            assert out.nlj == oo.nlj
        assert abs(dE) < etol * abs(Etot)
        scale = max(abs(oo.Pvir[k]) for k in range(3))
        for i in range(6):
            assert abs(out.Pvir[i] - oo.Pvir[i]) < 10 * etol * scale


def test_full_size_brine_against_the_oracle(c2):
    """one force call of the 194,304-site brine, both precisions, against
    the oracle (its FP64 CPU restatement, 16 threads) on the same input:
    forces, census, potential energy and tensor"""
    import oracle
    wl, bs, _ = c2
    r = np.mod(wl["r"], wl["L"])
    bso = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05,
                       provider=oracle.provider())
    _vs_oracle(bs, bso, r, "C2")


def test_full_size_drude_brine_against_the_oracle(c3):
    """the same for the Drude brine (97,200 sites + 32,600 Drudes) in its
    SCF state: the four charge terms of every pair at full size"""
    import oracle
    wl, bs, r = c3
    ns = bs.struct.ns
    r = r.copy()
    r[:ns] = np.mod(r[:ns], wl["L"])
    bso = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05,
                       provider=oracle.provider())
    _vs_oracle(bs, bso, r, "C3", polar=True)
