"""Force matching on the engine: the gradient of the force-matching loss
with respect to per-type-pair LJ sigma/epsilon (adjoint kernel) and the
type charges, against central finite differences of the loss (FP64) and
between the mixed and FP64 engines; the energy gradient against finite
differences of the energy."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.fit import ForceMatch

pytestmark = pytest.mark.gpu

PAIRS = [("O", "O"), ("NA", "O"), ("CL", "O"), ("NA", "CL")]
# q_H = -q_O/2, q_O, q_Na, q_Cl
QMAP = (["H", "O", "NA", "CL"], [0, 0, 0, 0],
        [[-0.5, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])


def setup(mode):
    d, meta = oracle.load_fixture("brine506")
    rng = np.random.default_rng(5)
    frames = [d["r"], d["r"] + rng.normal(scale=0.03, size=d["r"].shape)]
    L = meta["L"]
    kw = dict(epsr=meta["epsr"], epsk=meta["epsk"])
    ref = ForceMatch(meta["forcefield"], meta["counts"], L, meta["cutoff"], frames,
                     [np.zeros_like(f) for f in frames], PAIRS, QMAP, mode=abi.MDX_FP64, **kw)
    th0 = ref.theta0()
    tht = th0 * np.array([1.05] * 8 + [1, 1, 1]) + np.array([0] * 8 + [-0.02, 0.02, -0.02])
    targets = [f for f, _ in ref.forces(tht)]
    alpha, kappa = ref.bs.info["alpha"], ref.bs.info["kappa"]
    ref.close()
    fm = ForceMatch(meta["forcefield"], meta["counts"], L, meta["cutoff"], frames, targets,
                    PAIRS, QMAP, mode=mode, alpha=alpha, kappa=kappa, **kw)
    return fm, th0, tht


def test_force_matching_gradient_vs_finite_differences():
    fm, th0, tht = setup(abi.MDX_FP64)
    L, g = fm.loss_and_grad(th0)
    assert fm.loss(tht) < 1e-12 * L          # the targets are exactly reachable
    gfd = np.zeros_like(g)
    for k in range(len(th0)):
        h = 1e-4 * abs(th0[k])
        e = np.zeros_like(th0)
        e[k] = h
        gfd[k] = (fm.loss(th0 + e) - fm.loss(th0 - e)) / (2 * h)
    err = np.abs(g - gfd) / np.abs(gfd).max()
    print("loss %.6e  grad %s\n  fd %s\n  max rel err %.2e" % (L, g, gfd, err.max()), flush=True)
    assert err.max() < 1e-5
    # each component on its own scale too (sigma, epsilon, charges)
    assert np.all(np.abs(g - gfd) <= 1e-4 * np.abs(gfd) + 1e-6 * np.abs(gfd).max())
    fm.close()


def test_mixed_gradient_vs_fp64():
    fm64, th0, _ = setup(abi.MDX_FP64)
    fm32, _, _ = setup(abi.MDX_MIXED)
    L64, g64 = fm64.loss_and_grad(th0)
    L32, g32 = fm32.loss_and_grad(th0)
    e = np.abs(g32 - g64).max() / np.abs(g64).max()
    print("loss fp64 %.8e mixed %.8e, grad rel diff %.2e" % (L64, L32, e), flush=True)
    assert abs(L32 - L64) < 1e-4 * L64
    assert e < 1e-3
    fm64.close()
    fm32.close()


def test_energy_gradient_vs_finite_differences():
    fm, th0, _ = setup(abi.MDX_FP64)
    E, g = fm.energy_and_grad(th0)
    gfd = np.zeros_like(g)
    for k in range(len(th0)):
        h = 1e-4 * abs(th0[k])
        e = np.zeros_like(th0)
        e[k] = h
        Ep = sum(o.En_pot + o.En_el for _, o in fm.forces(th0 + e, measure=1))
        Em = sum(o.En_pot + o.En_el for _, o in fm.forces(th0 - e, measure=1))
        gfd[k] = (Ep - Em) / (2 * h)
    err = np.abs(g - gfd) / np.abs(gfd).max()
    print("E %.6e grad %s\n fd %s\n max rel err %.2e" % (E, g, gfd, err.max()), flush=True)
    assert err.max() < 1e-6
    fm.close()
