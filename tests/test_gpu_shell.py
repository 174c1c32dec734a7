"""The POLAR&1 shell-core repulsion on the GPU against the oracle: the
shell-core test system (cook's polar306 configuration in the Buckingham
form, shells on O and Cl, rep on O, Na and Cl), forces on centres and
Drudes, energies, virial and tensor, census; both precisions, the cell and
all-pair paths; one selffield iteration; kappa = 0 equals the system
without shell data."""
import numpy as np
import pytest

import oracle
import shellcore
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu

TOL = {abi.MDX_FP64: 1e-10, abi.MDX_MIXED: 1e-5}
# energies relative to |En_pot + En_el|: the measured calls of a Drude
# system run the FP64 pair kernel in both modes, the k-space stays FP32
ETOL = {abi.MDX_FP64: 1e-10, abi.MDX_MIXED: 2e-6}


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.mark.parametrize("w", [None, 0])     # default cell stencil, all-pair
@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_shell_forces_vs_oracle(mode, w):
    r, bs, _ = shellcore.system(provider_=engine_provider())
    ro, bso, _ = shellcore.system()
    print("shell-core polar306 mode %d stencil %s" % (mode, w), flush=True)
    f_ref, o_ref = oracle.forces(bso, ro, 1)
    e = Engine(bs, mode)
    if w is not None:
        e.set_stencil(w)
    f0, _ = e.forces(r, 0)
    f, out = e.forces(r, 1)
    e.close()
    err0, err1 = rel_rms(f0, f_ref), rel_rms(f, f_ref)
    print("  rel RMS force err unmeasured %.3e measured %.3e" % (err0, err1), flush=True)
    assert err0 < TOL[mode] and err1 < TOL[mode]
    assert out.npair == o_ref.npair
    Etot = abs(o_ref.En_pot + o_ref.En_el)
    for k in ("En_pot", "En_vir", "En_el"):
        dk = getattr(out, k) - getattr(o_ref, k)
        print("  %s %.10e diff %.2e" % (k, getattr(o_ref, k), dk), flush=True)
        assert abs(dk) < ETOL[mode] * Etot, k
    scale = max(abs(o_ref.Pvir[i]) for i in range(3))
    for i in range(6):
        assert abs(out.Pvir[i] - o_ref.Pvir[i]) < 10 * ETOL[mode] * scale


def test_shell_selffield_vs_oracle():
    # the Drude moves take the shell-core forces on the Drudes
    r, bs, _ = shellcore.system(provider_=engine_provider())
    _, bso, _ = shellcore.system()
    ns = bs.struct.ns
    omega = 0.6
    ro = r.copy()
    fo, oo = oracle.forces(bso, ro, 1)
    eso, _ = oracle.selffield(bso, ro, fo, oo, omega)
    e = Engine(bs, abi.MDX_FP64)
    rg = r.copy()
    f, out = e.forces(rg, 1)
    esg, _ = e.selffield(rg, f, out, omega)
    e.close()
    d = np.abs(rg[ns:] - ro[ns:]).max()
    print("selffield: Drude offsets engine vs oracle max diff %.2e, Eself %.10e / %.10e" % (d, esg, eso),
          flush=True)
    assert d < 1e-10 * np.abs(ro[ns:]).max()
    assert abs(esg - eso) < 1e-10 * abs(eso)


def test_shell_kappa_zero_is_no_shell():
    r, bs, meta = shellcore.system(kappa=0.0, provider_=engine_provider())
    bsn = oracle.fixture_system(dict(meta, forcefield=shellcore.buck_forcefield(meta["forcefield"])),
                                provider_=engine_provider())
    e, en = Engine(bs, abi.MDX_FP64), Engine(bsn, abi.MDX_FP64)
    f, _ = e.forces(r, 1)
    fn, _ = en.forces(r, 1)
    e.close()
    en.close()
    np.testing.assert_array_equal(f, fn)
