"""The exp-6 form (sim/exp6) on the GPU against the oracle, through the
C ABI: forces, energies, virial, tensor and census of a liquid brine (the
positions of cook's brine506 run, exp-6 parameters of the same minima),
both precisions, cell and all-pair paths; MD steps; the full-size brine
mixed against FP64; the LJ-only entry points refuse exp-6."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Batch, Engine
from macsimus_amd.system import build_system, engine_provider

pytestmark = pytest.mark.gpu

TOL = {abi.MDX_FP64: 1e-10, abi.MDX_MIXED: 1e-5}
ETOL = {abi.MDX_FP64: 1e-10, abi.MDX_MIXED: 2e-6}


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def brine():
    d, meta = oracle.load_fixture("brine506")
    ff = dict(meta["forcefield"])
    ff["form"] = "exp6"
    ff["type_parm"] = [workloads.EXP6_A[t] for t in ff["types"]]
    bs = oracle.fixture_system(dict(meta, forcefield=ff), provider_=engine_provider())
    bso = oracle.fixture_system(dict(meta, forcefield=ff))
    f_ref, o_ref = oracle.forces(bso, d["r"], 1)
    return d, meta, bs, bso, f_ref, o_ref


@pytest.mark.parametrize("w", [2, 0])
@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_exp6_forces_vs_oracle(brine, mode, w):
    d, meta, bs, bso, f_ref, o_ref = brine
    print("exp-6 brine506 mode %d stencil %d" % (mode, w), flush=True)
    e = Engine(bs, mode)
    e.set_stencil(w)
    f, out = e.forces(d["r"], 1)
    f0, _ = e.forces(d["r"], 0)
    e.close()
    err = rel_rms(f, f_ref)
    print("  rel RMS force err %.3e" % err, flush=True)
    assert err < TOL[mode]
    assert rel_rms(f0, f_ref) < TOL[mode]
    assert out.npair == o_ref.npair and out.nlj == o_ref.nlj
    # energies against the potential energy En.pot + En.el (the site-site
    # part alone is a small net sum of large terms of both signs)
    Etot = o_ref.En_pot + o_ref.En_el
    for k in ("En_pot", "En_vir", "En_el"):
        dk = getattr(out, k) - getattr(o_ref, k)
        print("  %s %.10e diff %.3e (%.2e of Epot)" % (k, getattr(o_ref, k), dk, dk / Etot), flush=True)
        assert abs(dk) < ETOL[mode] * abs(Etot), k
    scale = max(abs(o_ref.Pvir[i]) for i in range(3))
    for i in range(6):
        assert abs(out.Pvir[i] - o_ref.Pvir[i]) < 10 * ETOL[mode] * scale


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-9), (abi.MDX_MIXED, 1e-5)])
def test_exp6_md_steps_vs_oracle(brine, mode, tol):
    d, meta, bs, bso, _, _ = brine
    r = d["r"].copy()
    r1 = np.zeros_like(r)
    p = abi.Step(h=0.002, eps=1e-10, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=1.0,
                 measure_every=0)
    e = Engine(bs, mode)
    e.upload_state(r, r1)
    so = e.md_steps(p, 5)
    rg, _ = e.download_state()
    e.close()
    ro, r1o = r.copy(), r1.copy()
    for k in range(5):
        oo = oracle.md_step(bso, ro, r1o, p, measure=int(k == 4))
    dev = np.abs(rg - ro).max()
    print("exp-6 5 steps mode %d: max deviation %.3e A" % (mode, dev), flush=True)
    assert dev < tol
    Etot = oo.frc.En_pot + oo.frc.En_el
    assert abs(so.frc.En_pot + so.frc.En_el - Etot) < ETOL[mode] * abs(Etot)
    assert abs(so.frc.En_pot - oo.frc.En_pot) < ETOL[mode] * abs(Etot)


def test_exp6_full_size_mixed_vs_fp64():
    wl = workloads.brine_c2_exp6()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    res = {}
    for mode in (abi.MDX_FP64, abi.MDX_MIXED):
        print("C2 exp-6 mode %d" % mode, flush=True)
        e = Engine(bs, mode)
        res[mode] = e.forces(wl["r"], 1)
        e.close()
    (f64, o64), (f32, o32) = res[abi.MDX_FP64], res[abi.MDX_MIXED]
    err = rel_rms(f32, f64)
    print("  rel RMS force err mixed vs FP64 %.3e" % err, flush=True)
    assert err < 1e-5
    assert o32.npair == o64.npair and o32.nlj == o64.nlj
    Etot = o64.En_pot + o64.En_el
    print("  En_pot %.10e diff %.3e" % (o64.En_pot, o32.En_pot - o64.En_pot), flush=True)
    assert abs(o32.En_pot - o64.En_pot) < 2e-6 * abs(Etot)
    fmax = np.abs(f64).max()
    assert np.abs(f64.sum(0)).max() < 1e-9 * fmax * np.sqrt(len(f64))


def test_lj_only_entry_points_refuse_exp6(brine):
    """the LJ adjoint refuses an exp-6 context; the replica batch takes
    exp-6 systems (tests/test_gpu_batch.py) but, as for LJ, atomic ones
    only: the molecular brine is refused for that"""
    d, meta, bs, _, _, _ = brine
    e = Engine(bs, abi.MDX_FP64)
    with pytest.raises(RuntimeError, match="exp-6"):
        e.adjoint_lj(d["r"], np.ones_like(d["r"]))
    e.close()
    with pytest.raises(RuntimeError, match="atomic"):
        Batch([bs], abi.MDX_FP64)


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_buckingham_forces_vs_oracle(brine, mode):
    """the Buckingham form on the exp-6 kernels: each brine site as the
    Buckingham site of its exp-6 parameters (A, B, C of the identical-site
    mapping), geometric rules (comb_rule 5)"""
    import math
    d, meta, _, _, _, _ = brine
    ff = dict(meta["forcefield"], form="buck", comb_rule=5)
    A, B, Cc = [], [], []
    for t, E, R in zip(ff["types"], ff["type_EvdW_kcal"], ff["type_RvdW"]):
        a, sig, eps = workloads.EXP6_A[t], 2 * R, -E
        A.append(6 * eps * math.exp(a) / (a - 6))
        B.append(a / sig)
        Cc.append(eps * a * sig ** 6 / (a - 6))
    ff.update(type_EvdW_kcal=A, type_RvdW=B, type_parm=Cc)
    bs = oracle.fixture_system(dict(meta, forcefield=ff), provider_=engine_provider())
    bso = oracle.fixture_system(dict(meta, forcefield=ff))
    f_ref, o_ref = oracle.forces(bso, d["r"], 1)
    e = Engine(bs, mode)
    f, out = e.forces(d["r"], 1)
    e.close()
    err = rel_rms(f, f_ref)
    print("Buckingham brine506 mode %d: rel RMS force err %.3e" % (mode, err), flush=True)
    assert err < TOL[mode]
    assert out.npair == o_ref.npair and out.nlj == o_ref.nlj > 0
    Etot = o_ref.En_pot + o_ref.En_el
    assert abs(out.En_pot - o_ref.En_pot) < ETOL[mode] * abs(Etot)


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_busing_forces_vs_oracle(brine, mode):
    """the Busing form on the exp-6 kernels: the Buckingham sites of the
    brine's exp-6 parameters converted to Busing sites (rho = 1/2B,
    R = rho ln(A/2 f rho), C_i = sqrt(C); a site without repulsion gets
    EvdW < 0); against the oracle and against the Buckingham system of
    additive 1/B and the Busing rule for A (comb_rule 15) on the GPU"""
    import math
    from macsimus_amd.system import EUNIT_KCAL
    d, meta, _, _, _, _ = brine
    fb = 0.05 * 408.779826922048 ** 2 / EUNIT_KCAL           # f in kcal/mol/A
    base = meta["forcefield"]
    A, B, Cc, E, R, Ci = [], [], [], [], [], []
    for t, Ev, Rv in zip(base["types"], base["type_EvdW_kcal"], base["type_RvdW"]):
        a, sig, eps = workloads.EXP6_A[t], 2 * Rv, -Ev
        A.append(6 * eps * math.exp(a) / (a - 6))
        B.append(a / sig)
        Cc.append(eps * a * sig ** 6 / (a - 6))
        rho = 0.5 / B[-1]
        E.append(fb * rho if A[-1] > 0 else -1.0)
        R.append(rho * math.log(A[-1] / (2 * fb * rho)) if A[-1] > 0 else 0.0)
        Ci.append(math.sqrt(Cc[-1]))
    ffb = dict(base, form="busing", comb_rule=0, type_EvdW_kcal=E, type_RvdW=R, type_parm=Ci)
    ffk = dict(base, form="buck", comb_rule=15, type_EvdW_kcal=A, type_RvdW=B, type_parm=Cc)
    bso = oracle.fixture_system(dict(meta, forcefield=ffb))
    f_ref, o_ref = oracle.forces(bso, d["r"], 1)
    res = []
    for ff in (ffb, ffk):
        print("%s brine506 mode %d" % (ff["form"], mode), flush=True)
        e = Engine(oracle.fixture_system(dict(meta, forcefield=ff), provider_=engine_provider()),
                   mode)
        res.append(e.forces(d["r"], 1))
        e.close()
    (f, out), (fk, outk) = res
    err = rel_rms(f, f_ref)
    print("Busing brine506 mode %d: rel RMS force err %.3e" % (mode, err), flush=True)
    assert err < TOL[mode]
    assert out.npair == o_ref.npair and out.nlj == o_ref.nlj > 0
    Etot = o_ref.En_pot + o_ref.En_el
    assert abs(out.En_pot - o_ref.En_pot) < ETOL[mode] * abs(Etot)
    assert abs(out.En_vir - o_ref.En_vir) < ETOL[mode] * abs(o_ref.En_vir)
    # the same tables up to rounding: the same forces
    assert rel_rms(f, fk) < TOL[mode]
    assert out.En_pot == pytest.approx(outk.En_pot, rel=ETOL[mode])
