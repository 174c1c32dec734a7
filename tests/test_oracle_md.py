"""The oracle's Shake() step (SHAKE + leap-frog) against the NVE trajectory
that the reference cook dumped (tests/golden/spce216_nve.*)."""
import numpy as np

import oracle
from macsimus_amd import abi


def test_nve_trajectory_matches_reference():
    d, meta = oracle.load_fixture("spce216_nve")
    bs = oracle.fixture_system(meta)
    R = d["r"]                       # R[k] = r(t_{k-1}) for k >= 1
    r = R[2].copy()
    r1 = R[2] - R[1]
    p = abi.Step(h=0.002, eps=1e-12, maxit=0, thermostat=0, T=298, tau=0.1, nf=1, measure_every=1)
    for k in range(3, len(R)):
        out = oracle.md_step(bs, r, r1, p, measure=1)
        err = np.abs(r - R[k]).max()
        assert err < 2e-10, (k, err)
        assert out.shake_iter_max > 0


def test_constraints_hold_after_steps():
    d, meta = oracle.load_fixture("spce216_nve")
    bs = oracle.fixture_system(meta)
    R = d["r"]
    r = R[2].copy()
    r1 = R[2] - R[1]
    p = abi.Step(h=0.002, eps=1e-10, maxit=0, thermostat=0, T=298, tau=0.1, nf=1, measure_every=1)
    for _ in range(5):
        oracle.md_step(bs, r, r1, p)
    cij = bs.info["cons_ij"]
    dr = r[cij[:, 0]] - r[cij[:, 1]]
    rel = np.abs((dr ** 2).sum(1) / bs.info["cons_bondq"] - 1)
    assert rel.max() < 1e-9
