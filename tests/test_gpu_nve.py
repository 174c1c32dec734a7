"""Energy conservation of the device-resident NVE step (leap-frog + SHAKE,
no thermostat) over 10,000 steps of 2 fs on the 1000-SPC/E configuration
of the reference's MD run, both precisions: the drift of the total energy
(linear fit after the first 2 ps) and its fluctuation per degree of
freedom stay small, and the mixed engine's drift is not worse than the
FP64 path's by more than the noise of the estimate."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu

NBLK, PER, H = 100, 100, 0.002


def series(mode):
    d, meta = oracle.load_fixture("spce1000_md")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    nf = workloads.degrees_of_freedom(meta["forcefield"], meta["counts"])
    eng = Engine(bs, mode)
    eng.upload_state(d["r"], np.zeros_like(d["r"]))
    p = abi.Step(h=H, eps=1e-10, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=PER)
    E, K = [], []
    for _ in range(NBLK):
        so = eng.md_steps(p, PER)
        E.append(so.frc.En_pot + so.frc.En_el + so.En_kin)
        K.append(so.En_kin)
    eng.close()
    t = np.arange(1, NBLK + 1) * PER * H        # ps
    E = np.array(E)
    slope = np.polyfit(t[10:], E[10:], 1)[0] * 1000 / nf    # K per ns per dof
    return slope, E[10:].std() / nf, 2 * K[-1] / nf


def test_nve_drift_10k_steps():
    s64, sd64, T64 = series(abi.MDX_FP64)
    s32, sd32, T32 = series(abi.MDX_MIXED)
    print("NVE 10k steps: drift FP64 %.3f mixed %.3f K/ns/dof; std %.3f / %.3f K/dof; T %.1f / %.1f K"
          % (s64, s32, sd64, sd32, T64, T32), flush=True)
    assert 200 < T64 < 600 and 200 < T32 < 600      # the system heated up from rest
    for s, sd in ((s64, sd64), (s32, sd32)):
        assert abs(s) < 1.5
        assert sd < 0.1
    assert abs(s32) < abs(s64) + 1.0
