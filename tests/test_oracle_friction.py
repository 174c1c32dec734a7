"""The decoupled friction thermostats of cook's Shake() (sim/thermo.c:
107-135) in the oracle, pinned by cook: tests/golden/spce216_friction
holds the configuration of 5 Berendsen steps (cfg0) and the configuration
cook saved after 6 more steps from it with thermostat="BerendsenCM"
(friction on the molecular translations, 11), "BerendsenIN" (on the
intramolecular motion, 12) and "frictions" (both, T_tr_in = 1.3, 13),
tau.T = 0.05.  The oracle started from cfg0 ends on each of them."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, cookio, workloads

THERMOSTATS = {"tr": 11, "in": 12, "fr": 13}


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    d, meta = oracle.load_fixture("spce216_friction")
    td = tmp_path_factory.mktemp("fric")
    cfgs = {}
    for nm in ("cfg0", "cfg_tr", "cfg_in", "cfg_fr"):
        p = td / (nm + ".cfg")
        p.write_bytes(d[nm].tobytes())
        cfgs[nm] = cookio.read_cfg(p)
    return d, meta, cfgs


def step_params(meta, thermostat):
    nf = workloads.degrees_of_freedom(meta["forcefield"], meta["counts"])
    return abi.Step(h=meta["h"], eps=1e-12, maxit=0, thermostat=thermostat, T=meta["T"],
                    tau=meta["tauT"], nf=float(nf), measure_every=1,
                    nf_tr=3.0 * sum(meta["counts"]) - 3, T_tr_in=meta["T_tr_in"])


@pytest.mark.parametrize("tag", ["tr", "in", "fr"])
def test_oracle_ends_on_cooks_configuration(case, tag):
    d, meta, cfgs = case
    bs = oracle.fixture_system(meta)
    c0, c1 = cfgs["cfg0"], cfgs["cfg_" + tag]
    assert c1["thermostat"] == THERMOSTATS[tag]
    r = c0["cfgs"][0]["rp"].copy()
    r1 = c0["cfgs"][1]["rp"].copy()
    p = step_params(meta, THERMOSTATS[tag])
    for _ in range(meta["nsteps"]):
        out = oracle.md_step(bs, r, r1, p)
    dr = np.abs(r - c1["cfgs"][0]["rp"]).max()
    dr1 = np.abs(r1 - c1["cfgs"][1]["rp"]).max()
    print("%s: r %.2e r1 %.2e  kin_tr/kin %.3f" % (tag, dr, dr1, out.En_kin_tr / out.En_kin))
    assert dr < 1e-10 and dr1 < 1e-11
    assert 0 < out.En_kin_tr < out.En_kin


def test_thermostats_differ(case):
    """each run leaves the others' end points: the three branches are
    distinguished by cook's data"""
    d, meta, cfgs = case
    rs = [cfgs["cfg_" + t]["cfgs"][1]["rp"] for t in ("tr", "in", "fr")]
    for a in range(3):
        for b in range(a):
            assert np.abs(rs[a] - rs[b]).max() > 1e-6
