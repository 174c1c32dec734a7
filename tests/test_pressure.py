"""Pressure assembly (measurePconstraints, sim/constrd.c:115-265) and the
LJ cutoff correction (cutcorr, sim/siminit.c:665-722) of the C ABI,
against what the reference cook printed for the golden configurations."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.system import PUNIT, engine_provider

CASES = ["spce1000_lattice", "spce1000_md", "brine506", "polar306"]


@pytest.mark.parametrize("name", CASES)
def test_cutoff_correction_matches_reference(name):
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    e = abi.cutcorr(abi.load_engine(), bs)
    # cook prints En.corr with 10 significant digits
    assert abs(e - meta["Ecorr_KAA3"]) < 5e-9 * abs(meta["Ecorr_KAA3"])


@pytest.mark.parametrize("name", ["spce1000_lattice", "spce1000_md", "brine506"])
def test_virial_and_trace_pressure_match_reference(name):
    """Ewaldtest prints the virial pressure (En.vir+En.el)/3V and the
    pressure tensor of one force call; a force call alone is a step with
    no kinetic or constraint terms"""
    d, meta = oracle.load_fixture(name)
    bso = oracle.fixture_system(meta)
    f, out = oracle.forces(bso, d["r"])
    so = abi.StepOut()
    so.frc = out
    L = [meta["L"]] * 3
    p = abi.pressure(abi.load_engine(), so, L)
    assert abs(p.Pevir * PUNIT - meta["P_virial_Pa"]) < 1e-10 * abs(meta["P_virial_Pa"])
    assert abs(p.Ptr * PUNIT - meta["P_trace_Pa"]) < 1e-10 * abs(meta["P_trace_Pa"])
    # printed tensor: diagonal /3 (VO), off-diagonal raw (x V)
    V = meta["L"] ** 3
    ref = meta["Ptens_printed_Pa"]
    for i in range(3):
        assert abs(p.Ptens[i] * PUNIT / 3 - ref[i]) < 1e-10 * abs(ref[i])
    scale = max(abs(x) for x in ref[:3])
    for i in range(3, 6):
        assert abs(p.Ptens[i] * V * PUNIT - ref[i]) < 1e-9 * scale * 3 * V
    # the cutoff correction enters as Ecorr/V^2
    pc = abi.pressure(abi.load_engine(), so, L, Ecorr=meta["Ecorr_KAA3"])
    assert abs((pc.Ptr_c - pc.Ptr) - meta["Ecorr_KAA3"] / V ** 2) < 1e-12 * abs(pc.Ptr)


def test_kinetic_tensor_trace_is_twice_the_kinetic_energy():
    """the reference checks SUM(Pkin)/2 - Ekin = 0 (sim/constrd.c:157); with
    Pkinq the kinetic pressure scales, the trace identity keeps its form"""
    d, meta = oracle.load_fixture("spce216_nve")
    bs = oracle.fixture_system(meta)
    r = np.ascontiguousarray(d["r"][1])
    r1 = np.ascontiguousarray(d["r"][1] - d["r"][0])
    p = abi.Step(h=meta["h"], eps=1e-10, thermostat=0, T=298.0, tau=0.1, nf=1.0, measure_every=1)
    so = oracle.md_step(bs, r, r1, p)
    V = meta["L"] ** 3
    for q in (1.0, 0.99):
        pr = abi.pressure(abi.load_engine(), so, [meta["L"]] * 3, Pkinq=q)
        assert abs(sum(pr.Pkin[:3]) * V / q - 2 * so.En_kin) < 1e-9 * abs(so.En_kin)
        assert pr.Ptens[0] == pr.Pvir[0] + pr.Pkin[0]
    assert so.En_kin > 0 and so.virc != 0
