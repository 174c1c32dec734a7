"""Closed-form checks of the oracle's Ewald sum that do not rely on cook.

- Rock salt: the electrostatic energy of a perfect NaCl crystal is
  -M q^2/r0 per ion pair with the Madelung constant M = 1.747565 (the
  textbook value); the forces vanish by symmetry.
- Coulomb virial theorem: a converged Ewald sum is invariant under
  alpha, so scaling all lengths by lambda scales the energy by 1/lambda
  and the trace of the virial tensor equals the electrostatic energy
  (self term included); a dropped term or a wrong sign in the k-space or
  r-space tensor breaks it.
- Translation invariance of energy and forces; sum of forces zero.
Cook's own fixtures pin the oracle elsewhere; these pin it independently.
"""
import numpy as np

import oracle
from macsimus_amd import workloads
from macsimus_amd.system import ELECTRON, build_system

MADELUNG_NACL = 1.747565


def ionic_ff():
    """Na+ and Cl- without LJ (epsilon 0): the sum is pure Coulomb"""
    sp = lambda name, t, q: dict(name=name, site_type=[t], site_charge_e=[q], site_xyz=[[0, 0, 0]],
                                 site_excl=[[]], bonds=[])
    return dict(types=["NA", "CL"], type_mass=[22.98977, 35.453], type_EvdW_kcal=[0.0, 0.0],
                type_RvdW=[1.0, 1.0], species=[sp("NA", 0, 1.0), sp("CL", 1, -1.0)])


def rocksalt(n, r0):
    """n^3 sites of spacing r0 (n even), Na on even, Cl on odd parity,
    species-major order"""
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    even = g.sum(1) % 2 == 0
    return np.concatenate([g[even], g[~even]]).astype(float) * r0


def build_ionic(n, r0, cutoff):
    ff = ionic_ff()
    bs = build_system(ff, [n ** 3 // 2, n ** 3 // 2], [n * r0] * 3, cutoff, epsr=1e-7, epsk=1e-7,
                      provider=oracle.provider())
    return bs


def test_madelung_constant_of_rock_salt():
    n, r0 = 8, 2.8
    bs = build_ionic(n, r0, 10.0)
    r = rocksalt(n, r0) + 0.25 * r0              # off the box origin
    f, out = oracle.forces(bs, r, 1)
    npairs = n ** 3 // 2
    M = -out.En_el * r0 / (ELECTRON ** 2 * npairs)
    assert abs(M - MADELUNG_NACL) < 2e-6, M
    # forces vanish by symmetry (relative to q^2/r0^2)
    assert np.abs(f).max() < 1e-6 * ELECTRON ** 2 / r0 ** 2
    # cubic symmetry of the tensor: equal diagonal, zero off-diagonal
    pv = np.array(out.Pvir)
    assert np.ptp(pv[:3]) < 1e-8 * abs(pv[0]) and np.abs(pv[3:]).max() < 1e-8 * abs(pv[0])


def test_coulomb_virial_trace_equals_energy():
    """disordered melt: the 512 ions of the crystal displaced at random
    (no pair closer than 1.4 A), Coulomb only"""
    n, r0 = 8, 2.8
    bs = build_ionic(n, r0, 10.0)
    rng = np.random.default_rng(7)
    r = rocksalt(n, r0) + rng.uniform(-0.5, 0.5, size=(n ** 3, 3))
    f, out = oracle.forces(bs, r, 1)
    tr = sum(out.Pvir[:3])
    assert abs(tr - out.En_el) < 2e-6 * abs(out.En_el), (tr, out.En_el)
    # the virial theorem holds because the k-space tensor, not just its
    # trace, is consistent: a flipped off-diagonal sign breaks symmetry of
    # W under the exchange of axes for a rotated copy
    rot = r[:, [1, 2, 0]]
    _, out2 = oracle.forces(bs, rot, 1)
    pv, pv2 = np.array(out.Pvir), np.array(out2.Pvir)
    # xx,yy,zz,yz,zx,xy of the rotated copy = yy,zz,xx,zx,xy,yz of the original
    assert np.allclose(pv2, pv[[1, 2, 0, 4, 5, 3]], rtol=1e-9, atol=1e-9 * np.abs(pv).max())
    assert abs(out2.En_el - out.En_el) < 1e-10 * abs(out.En_el)


def test_translation_invariance_and_zero_net_force():
    sd = workloads.replicas_c4(nsys=4)[2]            # a salt melt, 1,604 atoms
    bs = workloads.build_replica(sd, provider=oracle.provider())
    f, out = oracle.forces(bs, sd["r"], 1)
    # shifted and wrapped back into the box (the oracle, like cook's
    # linked cells, takes positions inside it)
    shift = np.array([0.37, -1.91, 2.53]) * sd["L"]
    r2 = np.mod(sd["r"] + shift, sd["L"])
    f2, out2 = oracle.forces(bs, r2, 1)
    assert np.abs(f2 - f).max() < 1e-9 * np.abs(f).max()
    assert abs(out2.En_el - out.En_el) < 1e-10 * abs(out.En_el)
    assert abs(out2.En_pot - out.En_pot) < 1e-10 * abs(out.En_pot)
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).sum(1).max() * len(f) ** 0.5
