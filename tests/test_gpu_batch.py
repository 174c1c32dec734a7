"""Batched replicas (mdx_batch_*) against the oracle, system by system.

The batch holds atomic systems of the replica workload (binary salt melts
with Ewald, LJ mixtures without charges) of uneven size; every system's
forces, energies, virial, pressure tensor and pair census must equal what
the oracle (cook's forces(), cook/forces.c:93-224) gives for that system
alone, and the batched leap-frog + Berendsen steps what the oracle's
Shake() step (sim/constrd.c:622-1081) gives.  Bars: FP64 engine 1e-10
relative, mixed engine 1e-5 relative RMS force error, pair counts exact."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Batch
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu

_SYS = {}


def systems(n):
    """the first n replicas of the 256-system workload with their
    mdx_system_t (cached)"""
    if "all" not in _SYS:
        _SYS["all"] = workloads.replicas_c4()
    prov = engine_provider()
    out = []
    for k in range(n):
        if k not in _SYS:
            _SYS[k] = workloads.build_replica(_SYS["all"][k], provider=prov)
        out.append((_SYS["all"][k], _SYS[k]))
    return out


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def check_out(o, ref, tol, name, exact_nlj=True):
    assert o.npair == ref.npair, name
    if exact_nlj:
        assert o.nlj == ref.nlj, name
    for k in ("En_pot", "En_vir"):
        assert rel(getattr(o, k), getattr(ref, k)) < tol, (name, k, getattr(o, k), getattr(ref, k))
    # electrostatics: relative to the size of the k-space and r-space parts
    scale = max(abs(ref.En_el), abs(ref.En_kspace), 1e-300)
    assert abs(o.En_el - ref.En_el) < tol * scale, (name, o.En_el, ref.En_el)
    assert abs(o.En_kspace - ref.En_kspace) < tol * scale, (name, o.En_kspace, ref.En_kspace)
    pv, pr = np.array(o.Pvir), np.array(ref.Pvir)
    assert np.abs(pv - pr).max() < tol * np.abs(pr).max(), (name, pv, pr)


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-10), (abi.MDX_MIXED, 1e-5)])
def test_batch_forces_match_oracle(mode, tol):
    sy = systems(10)
    b = Batch([bs for _, bs in sy], mode)
    info = b.info()
    assert info["nsys"] == 10 and info["sites"] == sum(bs.struct.ns for _, bs in sy)
    print("batch", info, flush=True)
    rs = [sd["r"] for sd, _ in sy]
    fs, outs = b.forces(rs, measure=1)
    for (sd, bs), f, o in zip(sy, fs, outs):
        fo, oo = oracle.forces(bs, sd["r"], 1)
        err = rel_rms(f, fo)
        print("%s %s ns %d rel rms %.2e pairs %d" % (sd["name"], sd["kind"], bs.struct.ns, err,
                                                      o.npair), flush=True)
        assert err < tol, (sd["name"], err)
        check_out(o, oo, tol, sd["name"], exact_nlj=(mode == abi.MDX_FP64))
    # unmeasured call (the LJ force in its other algebraic form, cook/
    # ljqqnp.c:203-213): same forces to rounding, census still filled
    fs2, outs2 = b.forces(rs, measure=0)
    for f, f2, o, o2 in zip(fs, fs2, outs, outs2):
        assert rel_rms(f2, f) < (1e-12 if mode == abi.MDX_FP64 else tol)
        assert o2.npair == o.npair and o2.En_pot == 0
    b.close()


def test_batch_ragged_and_wrapped():
    """sites given outside the box (shifted by whole and half boxes) give
    the forces of the wrapped configuration; a batch of one system and a
    batch with a system repeated agree with the oracle"""
    sy = systems(4)
    sd, bs = sy[3]
    L = sd["L"]
    rng = np.random.default_rng(11)
    shifted = sd["r"] + L * rng.integers(-3, 4, size=sd["r"].shape)
    b = Batch([bs, sy[2][1], bs], abi.MDX_FP64)
    fs, outs = b.forces([shifted, sy[2][0]["r"], sd["r"]], 1)
    fo, oo = oracle.forces(bs, sd["r"], 1)
    assert rel_rms(fs[0], fo) < 1e-10 and rel_rms(fs[2], fo) < 1e-10
    assert outs[0].npair == outs[2].npair == oo.npair
    f2, o2 = oracle.forces(sy[2][1], sy[2][0]["r"], 1)
    assert rel_rms(fs[1], f2) < 1e-10
    b.close()
    b1 = Batch([bs], abi.MDX_MIXED)
    f1, o1 = b1.forces([sd["r"]], 1)
    assert rel_rms(f1[0], fo) < 1e-5 and o1[0].npair == oo.npair
    b1.close()


def test_batch_stencils_agree():
    """cells of cutoff/w for w = 1, 2, 3 give the same census and forces"""
    sy = systems(6)
    b = Batch([bs for _, bs in sy], abi.MDX_FP64)
    rs = [sd["r"] for sd, _ in sy]
    ref = None
    for w in (1, 2, 3):
        b.set_stencil(w)
        assert b.info()["stencil"] >= w
        fs, outs = b.forces(rs, 1)
        if ref is None:
            ref = (fs, outs)
            continue
        for f, f0, o, o0 in zip(fs, ref[0], outs, ref[1]):
            assert o.npair == o0.npair and o.nlj == o0.nlj
            assert rel_rms(f, f0) < 1e-12
    w, kzc, kparts, ms1, ms2 = b.autotune(rs)
    assert 1 <= w <= 3 and kzc in (8, 16, 24, 32) and 1 <= kparts <= 32 and 0 < ms2 <= ms1 * 1.5
    info = b.info()
    assert (info["stencil"], info["kchunk"], info["kparts"]) == (w, kzc, kparts)
    fs, outs = b.forces(rs, 1)
    for f, f0, o, o0 in zip(fs, ref[0], outs, ref[1]):
        assert o.npair == o0.npair and rel_rms(f, f0) < 1e-12
    b.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_batch_kspace_splits_agree(mode):
    """every kz chunk and work split of the k-space gives the same forces
    and energies (FP64: to rounding; mixed: to its bar)"""
    sy = systems(6)
    b = Batch([bs for _, bs in sy], mode)
    rs = [sd["r"] for sd, _ in sy]
    ref = b.forces(rs, 1)
    tol = 1e-12 if mode == abi.MDX_FP64 else 1e-6
    for mf in ((0, 1, 2) if mode == abi.MDX_MIXED else (0,)):
        if mode == abi.MDX_MIXED:
            b.set_kspace_mfma(mf)
        for kzc in (8, 16, 24, 32):
            for kparts in (1, 3, 32):
                b.set_kspace(kzc, kparts)
                assert b.info()["kmfma"] == mf
                fs, outs = b.forces(rs, 1)
                for f, f0, o, o0 in zip(fs, ref[0], outs, ref[1]):
                    assert rel_rms(f, f0) < tol, (mf, kzc, kparts)
                    assert abs(o.En_el - o0.En_el) <= tol * abs(o0.En_el) + 1e-12
    if mode == abi.MDX_FP64:
        with pytest.raises(RuntimeError):
            b.set_kspace_mfma(True)
    else:
        with pytest.raises(RuntimeError):
            b.set_kspace_mfma(3)
    b.close()


def make_step(sd, thermostat=1):
    nf = 3.0 * sum(sd["counts"]) - 3
    return abi.Step(h=0.005, eps=1e-8, maxit=0, thermostat=thermostat, T=sd["T"], tau=0.1,
                    nf=nf, measure_every=0)


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-9), (abi.MDX_MIXED, 1e-5)])
def test_batch_md_matches_oracle(mode, tol):
    """6 leap-frog + Berendsen steps of 6 systems (each its own T, nf)"""
    sy = systems(6)
    b = Batch([bs for _, bs in sy], mode)
    rs = [sd["r"].copy() for sd, _ in sy]
    r1s = [sd["v"] * 0.005 for sd, _ in sy]
    ps = [make_step(sd) for sd, _ in sy]
    b.upload_state(rs, r1s)
    nstep = 6
    outs = b.md_steps(ps, nstep)
    rg, r1g = b.download_state()
    for k, (sd, bs) in enumerate(sy):
        r, r1 = rs[k].copy(), r1s[k].copy()
        for s in range(nstep):
            oo = oracle.md_step(bs, r, r1, ps[k], measure=int(s == nstep - 1))
        L = sd["L"]
        dr = np.abs(rg[k] - r).max()
        dv = rel_rms(r1g[k], r1)
        print("%s md: |dr| %.2e rel dv %.2e Ekin %.10g / %.10g lambda %.12f / %.12f"
              % (sd["name"], dr, dv, outs[k].En_kin, oo.En_kin, outs[k].lambda_, oo.lambda_),
              flush=True)
        assert dr < tol * L and dv < tol * 10
        assert rel(outs[k].En_kin, oo.En_kin) < tol
        assert abs(outs[k].lambda_ - oo.lambda_) < tol
        assert outs[k].frc.npair == oo.frc.npair
        assert rel(outs[k].frc.En_pot, oo.frc.En_pot) < tol
        pk, pko = np.array(outs[k].Pkin), np.array(oo.Pkin)
        assert np.abs(pk - pko).max() < tol * np.abs(pko).max()
    b.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_batch_graph_replay_matches_eager(mode):
    """the whole-batch step replayed from its hipGraph gives the bits of
    eager launches, round after round, with measured steps, downloads and
    synchronizations between the replays; a parameter change re-captures"""
    sy = systems(8)
    rs = [sd["r"] for sd, _ in sy]
    r1s = [sd["v"] * 0.005 for sd, _ in sy]
    ps = [make_step(sd) for sd, _ in sy]
    bats = []
    for graph in (True, False):
        b = Batch([bs for _, bs in sy], mode)
        b.set_graph(graph)
        b.upload_state(rs, r1s)
        bats.append(b)
    for rnd in range(5):
        if rnd == 3:
            ps = [make_step(sd, thermostat=0) for sd, _ in sy]
        outs = [b.md_steps(ps, 7) for b in bats]
        sg, se = [b.download_state() for b in bats]
        for k in range(len(sy)):
            assert np.array_equal(sg[0][k], se[0][k]) and np.array_equal(sg[1][k], se[1][k]), rnd
            assert outs[0][k].En_kin == outs[1][k].En_kin
            assert outs[0][k].frc.En_el == outs[1][k].frc.En_el
    for b in bats:
        b.close()


def test_batch_rejects_molecular_systems():
    d, meta = oracle.load_fixture("brine506")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    with pytest.raises(RuntimeError):
        Batch([bs], abi.MDX_FP64)


def test_batch_full_256_systems():
    """the whole 256-system replica batch (828,442 atoms): mixed against
    FP64 for every system (1e-5 relative RMS force error, equal census),
    the three largest systems against the oracle, then 20 graph-replayed
    MD steps with finite energies and Berendsen factors near 1"""
    systems(0)
    allsys = _SYS["all"]
    prov = engine_provider()
    bss = [workloads.build_replica(sd, provider=prov) for sd in allsys]
    rs = [sd["r"] for sd in allsys]
    b64 = Batch(bss, abi.MDX_FP64)
    f64, o64 = b64.forces(rs, 1)
    b64.close()
    bm = Batch(bss, abi.MDX_MIXED)
    fm, om = bm.forces(rs, 1)
    worst = 0.0
    for k in range(len(allsys)):
        e = rel_rms(fm[k], f64[k])
        worst = max(worst, e)
        assert e < 1e-5, (k, e)
        assert om[k].npair == o64[k].npair
        assert rel(om[k].En_pot, o64[k].En_pot) < 1e-5
    print("256 systems: worst mixed-vs-FP64 rel rms %.2e, pairs %d"
          % (worst, sum(o.npair for o in o64)), flush=True)
    big = sorted(range(len(allsys)), key=lambda k: -bss[k].struct.ns)[:3]
    for k in big:
        fo, oo = oracle.forces(bss[k], rs[k], 1)
        e = rel_rms(f64[k], fo)
        print("%s ns %d FP64 vs oracle %.2e" % (allsys[k]["name"], bss[k].struct.ns, e), flush=True)
        assert e < 1e-10
        check_out(o64[k], oo, 1e-10, allsys[k]["name"])
    h = 0.005
    ps = [make_step(sd) for sd in allsys]
    bm.upload_state(rs, [sd["v"] * h for sd in allsys])
    outs = bm.md_steps(ps, 20)
    for k, o in enumerate(outs):
        assert np.isfinite(o.frc.En_pot) and np.isfinite(o.En_kin)
        assert 0.9 < o.lambda_ < 1.1, (k, o.lambda_)
    bm.close()


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-10), (abi.MDX_MIXED, 1e-5)])
def test_batch_exp6_matches_oracle(mode, tol):
    """replicas with the exp-6 form (a = 14 for every type, the LJ minima)
    against the oracle; a batch mixing forms is refused"""
    base = systems(6)
    prov = engine_provider()
    sy = []
    for sd, _ in base:
        ff = dict(sd["ff"], form="exp6", type_parm=[14.0] * len(sd["ff"]["types"]))
        sd6 = dict(sd, ff=ff)
        sy.append((sd6, workloads.build_replica(sd6, provider=prov),
                   workloads.build_replica(sd6, provider=oracle.provider())))
    b = Batch([bs for _, bs, _ in sy], mode)
    fs, outs = b.forces([sd["r"] for sd, _, _ in sy], measure=1)
    for (sd, bs, bso), f, o in zip(sy, fs, outs):
        fo, oo = oracle.forces(bso, sd["r"], 1)
        err = rel_rms(f, fo)
        print("exp-6 %s %s rel rms %.2e" % (sd["name"], sd["kind"], err), flush=True)
        assert err < tol, (sd["name"], err)
        check_out(o, oo, tol, sd["name"], exact_nlj=(mode == abi.MDX_FP64))
    b.close()
    with pytest.raises(RuntimeError, match="same site-site form"):
        Batch([sy[0][1], base[1][1]], mode)
