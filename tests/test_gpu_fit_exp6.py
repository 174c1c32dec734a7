"""Force matching with the exp-6 form: the gradient of the force-matching
loss with respect to per-type-pair exp-6 sigma (distance of the minimum),
epsilon and stiffness a (adjoint kernel mdx_adjoint_exp6) and the type
charges, against central finite differences of the loss (FP64); the
energy gradient against finite differences of the energy; mixed against
FP64."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.fit import ForceMatch

pytestmark = pytest.mark.gpu

PAIRS = [("O", "O"), ("NA", "O"), ("CL", "O")]
QMAP = (["H", "O", "NA", "CL"], [0, 0, 0, 0],
        [[-0.5, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
NPL = 3 * len(PAIRS)


def setup(mode):
    d, meta = oracle.load_fixture("brine506")
    ff = dict(meta["forcefield"], form="exp6")
    ff["type_parm"] = [workloads.EXP6_A[t] for t in ff["types"]]
    rng = np.random.default_rng(6)
    frames = [d["r"], d["r"] + rng.normal(scale=0.03, size=d["r"].shape)]
    L = meta["L"]
    kw = dict(epsr=meta["epsr"], epsk=meta["epsk"])
    ref = ForceMatch(ff, meta["counts"], L, meta["cutoff"], frames,
                     [np.zeros_like(f) for f in frames], PAIRS, QMAP, mode=abi.MDX_FP64, **kw)
    th0 = ref.theta0()
    assert len(th0) == NPL + 3
    tht = th0 * np.array([1.03] * NPL + [1, 1, 1]) + np.array([0] * NPL + [-0.02, 0.02, -0.02])
    targets = [f for f, _ in ref.forces(tht)]
    alpha, kappa = ref.bs.info["alpha"], ref.bs.info["kappa"]
    ref.close()
    fm = ForceMatch(ff, meta["counts"], L, meta["cutoff"], frames, targets, PAIRS, QMAP,
                    mode=mode, alpha=alpha, kappa=kappa, **kw)
    return fm, th0, tht


def fd(fun, th0):
    g = np.zeros_like(th0)
    for k in range(len(th0)):
        h = 1e-4 * abs(th0[k])
        e = np.zeros_like(th0)
        e[k] = h
        g[k] = (fun(th0 + e) - fun(th0 - e)) / (2 * h)
    return g


def test_exp6_force_matching_gradient_vs_finite_differences():
    fm, th0, tht = setup(abi.MDX_FP64)
    L, g = fm.loss_and_grad(th0)
    assert fm.loss(tht) < 1e-12 * L
    gfd = fd(fm.loss, th0)
    err = np.abs(g - gfd) / np.abs(gfd).max()
    print("loss %.6e\n  grad %s\n  fd   %s\n  max rel err %.2e" % (L, g, gfd, err.max()), flush=True)
    assert err.max() < 1e-5
    assert np.all(np.abs(g - gfd) <= 1e-4 * np.abs(gfd) + 1e-6 * np.abs(gfd).max())
    fm.close()


def test_exp6_energy_gradient_vs_finite_differences():
    fm, th0, _ = setup(abi.MDX_FP64)
    E, g = fm.energy_and_grad(th0)
    gfd = fd(lambda th: sum(o.En_pot + o.En_el for _, o in fm.forces(th, measure=1)), th0)
    err = np.abs(g - gfd) / np.abs(gfd).max()
    print("E %.6e\n  grad %s\n  fd   %s\n  max rel err %.2e" % (E, g, gfd, err.max()), flush=True)
    assert err.max() < 1e-6
    fm.close()


def test_exp6_mixed_gradient_vs_fp64():
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
