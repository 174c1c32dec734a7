"""Closed-form checks of the engines' Ewald sums on the GPU: the Madelung
constant of rock salt, the Coulomb virial theorem (trace of the virial
tensor = electrostatic energy) and translation invariance, for the
single-system engine in both modes and for the replica batch."""
import numpy as np
import pytest

import test_oracle_invariants as inv
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Batch, Engine
from macsimus_amd.system import ELECTRON, build_system

pytestmark = pytest.mark.gpu

TOL = {abi.MDX_FP64: 2e-6, abi.MDX_MIXED: 1e-5}


def ionic(n=8, r0=2.8):
    return build_system(inv.ionic_ff(), [n ** 3 // 2, n ** 3 // 2], [n * r0] * 3, 10.0,
                        epsr=1e-7, epsk=1e-7)


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_engine_madelung_and_virial(mode):
    n, r0 = 8, 2.8
    bs = ionic(n, r0)
    e = Engine(bs, mode)
    f, out = e.forces(inv.rocksalt(n, r0) + 0.25 * r0, 1)
    M = -out.En_el * r0 / (ELECTRON ** 2 * (n ** 3 // 2))
    print("mode %d: Madelung %.8f" % (mode, M), flush=True)
    assert abs(M - inv.MADELUNG_NACL) < TOL[mode]
    assert np.abs(f).max() < TOL[mode] * ELECTRON ** 2 / r0 ** 2
    rng = np.random.default_rng(7)
    r = inv.rocksalt(n, r0) + rng.uniform(-0.5, 0.5, size=(n ** 3, 3))
    f, out = e.forces(r, 1)
    tr = sum(out.Pvir[:3])
    assert abs(tr - out.En_el) < TOL[mode] * abs(out.En_el), (tr, out.En_el)
    # translation by a fraction of the box (no wrapping needed: the engine
    # wraps), same energy and forces
    f2, out2 = e.forces(r + np.array([0.31, -0.77, 1.43]) * n * r0, 1)
    assert np.abs(f2 - f).max() < TOL[mode] * np.abs(f).max()
    assert abs(out2.En_el - out.En_el) < TOL[mode] * abs(out.En_el)
    e.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_batch_madelung_and_virial(mode):
    """the rock-salt crystal and its disordered copy side by side with a
    salt-melt replica; each system's energy and tensor obey the same
    closed forms"""
    n, r0 = 8, 2.8
    bs = ionic(n, r0)
    sd = workloads.replicas_c4(nsys=4)[2]
    bsr = workloads.build_replica(sd)
    rng = np.random.default_rng(7)
    rd = inv.rocksalt(n, r0) + rng.uniform(-0.5, 0.5, size=(n ** 3, 3))
    b = Batch([bs, bsr, bs], mode)
    fs, outs = b.forces([inv.rocksalt(n, r0) + 0.25 * r0, sd["r"], rd], 1)
    M = -outs[0].En_el * r0 / (ELECTRON ** 2 * (n ** 3 // 2))
    assert abs(M - inv.MADELUNG_NACL) < TOL[mode], M
    tr = sum(outs[2].Pvir[:3])
    assert abs(tr - outs[2].En_el) < TOL[mode] * abs(outs[2].En_el)
    b.close()
