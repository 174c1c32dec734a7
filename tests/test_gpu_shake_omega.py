"""SHAKE sweep counts and relaxation omega per species on the device: the
drop-in for cook's omegac optimization (doconstrit(), sim/norm.c:32-78),
which stays on the caller's side.  The engine's sweep counts equal the
oracle's (sim/constrd.c:930), per species, up to a sweep or two: FP64
engine and oracle agree to ~1e-13, and a molecule whose largest relative
violation lands within that of eps after a sweep may take one sweep more
or less.  A new omega per species gives the oracle's steps with that
omega; and cook's optimization loop, restated here with a seeded generator
in place of cook's rnd() and fed the engine's counts, drives engine and
oracle through the same omega sequence to the same trajectory."""



def close(a, b):
    """sweep counts equal up to that rounding: 2 sweeps or 1e-4"""
    return abs(int(a) - int(b)) <= max(2, 1e-4 * max(a, b))
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu


def _start(name):
    d, meta = oracle.load_fixture(name)
    R = d["r"]
    return meta, R[-1].copy(), (R[-1] - R[-2]).copy()


def _step():
    return abi.Step(h=0.002, eps=1e-12, maxit=0, thermostat=0, T=298, tau=0.1, nf=1,
                    measure_every=1)


def test_sweep_count_equals_the_oracle():
    meta, r0, r10 = _start("spce216_nve")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    bso = oracle.fixture_system(meta)
    e = Engine(bs, abi.MDX_FP64)
    e.upload_state(r0, r10)
    r, r1 = r0.copy(), r10.copy()
    p = _step()
    for k in range(4):
        out = e.md_steps(p, 1)
        oo = oracle.md_step(bso, r, r1, p, measure=1)
        print("step %d: nit engine %d oracle %d" % (k, out.shake_nit, oo.shake_nit), flush=True)
        assert oo.shake_nit > 0 and close(out.shake_nit, oo.shake_nit)
        assert e.shake_nit()[0] == out.shake_nit
    # several steps in one call: the sum of the steps
    e.upload_state(r0, r10)
    out = e.md_steps(p, 3)
    r, r1 = r0.copy(), r10.copy()
    tot = sum(oracle.md_step(bso, r, r1, p, measure=1).shake_nit for _ in range(3))
    assert close(out.shake_nit, tot)
    e.close()


def test_species_and_omega():
    """brine506: ions (species 0, 1) have no constraints, water (2) does;
    omega per species for water gives the oracle's steps with it"""
    d, meta = oracle.load_fixture("brine506")
    counts = meta["counts"]
    msp = np.repeat(np.arange(len(counts)), counts).astype(np.int32)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    bso = oracle.fixture_system(meta)
    r0 = d["r"].copy()
    r10 = np.zeros_like(r0)
    p = _step()
    e = Engine(bs, abi.MDX_FP64)
    e.set_shake_species(msp, len(counts))
    e.set_shake_omega([0.6, 0.6, 0.9])
    for m in range(bso.struct.nmol):
        bso.struct.mol_omega[m] = 0.9 if msp[m] == 2 else 0.6
    e.upload_state(r0, r10)
    r, r1 = r0.copy(), r10.copy()
    for k in range(3):
        out = e.md_steps(p, 1)
        oo = oracle.md_step(bso, r, r1, p, measure=1)
        nit = e.shake_nit()
        assert nit[0] == nit[1] == 0 and nit[2] == out.shake_nit
        assert close(nit[2], oo.shake_nit)
    re_, r1e = e.download_state()
    assert np.abs(re_ - r).max() < 1e-12 and np.abs(r1e - r1).max() < 1e-12
    with pytest.raises(RuntimeError):
        e.set_shake_omega([0.6, 0.6, 2.5])
    with pytest.raises(RuntimeError):
        e.set_shake_species(np.full(len(msp), 7, np.int32), 3)
    e.close()


def _doconstrit(st, nit, rng):
    """the omegac optimization of doconstrit() (sim/norm.c:32-78) for one
    species, rng.random() in place of cook's rnd()"""
    if st["omegaclast"]:
        if nit and st["nitlast"]:
            if nit < st["nitlast"]:
                st["omegac"] = st["omegaclast"] * (1 - st["weight"]) + st["omegac"] * st["weight"]
            else:
                st["omegac"] = st["omegaclast"]
        st["omegaclast"] = 0
    elif rng.random() < st["often"]:
        st["omegaclast"] = st["omegac"]
        st["nitlast"] = nit
        st["omegac"] += rng.random() * st["d"]
        st["d"] = -st["d"]
    else:
        st["omegaclast"] = 0
    st["often"] = (1 - 1 / 200) * st["often"] + 0.1 / 200
    return st["omegac"] / 2


def test_cooks_omega_optimization_loop():
    """one step per call (cook runs doconstrit() every step), the sweep
    count to the optimizer, its omega back: engine (FP64) and oracle follow
    the same omega sequence to the same positions, and the sequence moves
    omega"""
    meta, r0, r10 = _start("spce216_nve")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    bso = oracle.fixture_system(meta)
    nmol = bso.struct.nmol
    p = _step()
    p.measure_every = 2
    e = Engine(bs, abi.MDX_FP64)
    e.upload_state(r0, r10)
    r, r1 = r0.copy(), r10.copy()
    st = dict(omegac=1.2, omegaclast=0, nitlast=0, d=0.03125, weight=0.25 / 4 * 4, often=1.0)
    rng = np.random.default_rng(7)
    oms = []
    for k in range(20):
        out = e.md_steps(p, 1)
        oo = oracle.md_step(bso, r, r1, p, measure=1)
        assert close(out.shake_nit, oo.shake_nit), k
        om = _doconstrit(st, out.shake_nit, rng)
        oms.append(om)
        e.set_shake_omega([om])
        for m in range(nmol):
            bso.struct.mol_omega[m] = om
    re_, _ = e.download_state()
    e.close()
    print("omega sequence", oms, flush=True)
    assert len(set(oms)) > 2
    assert np.abs(re_ - r).max() < 1e-11
