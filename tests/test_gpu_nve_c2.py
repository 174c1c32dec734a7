"""Energy conservation at full size: the 194,304-site brine (64,000 SPC/E +
1,152 NaCl, 12 A cutoff, Ewald at el.epsr = el.epsk = 0.05), 1,000 steps of
Berendsen at 298 K from the lattice in the mixed engine, then 10,000 NVE
steps of 2 fs (leap-frog + SHAKE, no thermostat) from that same state in
each precision.  The drift of the total energy (linear fit after the first
2 ps) and its fluctuation per degree of freedom stay within the bounds of
the 1000-SPC/E NVE test, and the mixed engine's drift is not worse than
the FP64 path's by more than the noise of the estimate."""
import numpy as np
import pytest

from macsimus_amd import abi, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import build_system

pytestmark = pytest.mark.gpu

NBLK, PER, H, EQUIL = 100, 100, 0.002, 1000
RESULTS = {}


@pytest.fixture(scope="module")
def c2():
    wl = workloads.brine_c2()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    nf = workloads.degrees_of_freedom(wl["ff"], wl["counts"])
    eng = Engine(bs, abi.MDX_MIXED)
    eng.upload_state(wl["r"], np.zeros_like(wl["r"]))
    p = abi.Step(h=H, eps=1e-10, maxit=0, thermostat=1, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=0)
    eng.md_steps(p, EQUIL)
    r, v = eng.download_state()
    eng.close()
    print("C2: %d sites equilibrated for %d steps" % (r.shape[0], EQUIL), flush=True)
    return bs, nf, r, v


def series(c2, mode):
    bs, nf, r, v = c2
    eng = Engine(bs, mode)
    eng.upload_state(r, v)
    p = abi.Step(h=H, eps=1e-10, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=PER)
    E, K = [], []
    for b in range(NBLK):
        so = eng.md_steps(p, PER)
        E.append(so.frc.En_pot + so.frc.En_el + so.En_kin)
        K.append(so.En_kin)
        if (b + 1) % 20 == 0:
            print("  %d steps, Etot %.6e" % ((b + 1) * PER, E[-1]), flush=True)
    eng.close()
    t = np.arange(1, NBLK + 1) * PER * H        # ps
    E = np.array(E)
    slope = np.polyfit(t[10:], E[10:], 1)[0] * 1000 / nf    # K per ns per dof
    return slope, E[10:].std() / nf, 2 * K[-1] / nf


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", [abi.MDX_MIXED, abi.MDX_FP64], ids=["mixed", "fp64"])
def test_c2_nve_drift_10k_steps(c2, mode):
    s, sd, T = series(c2, mode)
    RESULTS[mode] = s
    print("C2 NVE 10k steps (%s): drift %.3f K/ns/dof, std %.4f K/dof, T %.1f K"
          % ("mixed" if mode == abi.MDX_MIXED else "fp64", s, sd, T), flush=True)
    assert 200 < T < 600
    assert abs(s) < 1.5
    assert sd < 0.1


def test_c2_mixed_drift_not_worse_than_fp64():
    if len(RESULTS) < 2:
        pytest.skip("needs both NVE runs of this module")
    assert abs(RESULTS[abi.MDX_MIXED]) < abs(RESULTS[abi.MDX_FP64]) + 1.0
