"""Nose-Hoover with the time-reversible velocity predictor (TRVP) in the
oracle, pinned by cook: tests/golden/spce216_nose holds the configuration
cook saved after 5 Berendsen steps (cfg0: r, r1, logs), the positions of
the 8 force calls of the following run with thermostat="Nose", tau.T=0.05
and -p3 (TRVP length 3), and the configuration cook saved after it (cfg1:
r, r1 and the Nose variable logs, cfg[1]->logs).  The oracle started from
cfg0 follows the dumped positions step by step and ends at cfg1, positions
and the thermostat variable alike.  The configuration files themselves
round-trip through cookio.read_cfg / write_cfg byte for byte."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, cookio, workloads


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    d, meta = oracle.load_fixture("spce216_nose")
    td = tmp_path_factory.mktemp("nose")
    cfgs = []
    for nm in ("cfg0", "cfg1"):
        p = td / (nm + ".cfg")
        p.write_bytes(d[nm].tobytes())
        cfgs.append(cookio.read_cfg(p))
    return d, meta, cfgs, td


def test_cfg_files_round_trip(case):
    d, meta, cfgs, td = case
    for nm, c in zip(("cfg0", "cfg1"), cfgs):
        out = td / ("w_" + nm + ".cfg")
        cookio.write_cfg(out, c)
        assert out.read_bytes() == d[nm].tobytes()
        assert c["ns"] == d["r"].shape[1] and c["N"] == meta["counts"][0]
        assert c["h"] == meta["h"] and len(c["cfgs"]) == 2
    assert cfgs[0]["thermostat"] == 1 and cfgs[1]["thermostat"] == 2
    assert cfgs[1]["tauT"] == meta["tauT"]


def nose_step_params(meta):
    nf = workloads.degrees_of_freedom(meta["forcefield"], meta["counts"])
    return abi.Step(h=meta["h"], eps=1e-12, maxit=0, thermostat=2, T=meta["T"],
                    tau=meta["tauT"], nf=float(nf), measure_every=1, trvp_k=meta["pred"])


def test_oracle_follows_cook(case):
    d, meta, (c0, c1), _ = case
    bs = oracle.fixture_system(meta)
    r = c0["cfgs"][0]["rp"].copy()
    r1 = c0["cfgs"][1]["rp"].copy()
    nose = np.array([c0["cfgs"][0]["logs"], c0["cfgs"][1]["logs"], 0.0, 0.0])
    hist = np.zeros((meta["pred"], 1 + r.size))
    kk = 0
    p = nose_step_params(meta)
    for s in range(meta["nsteps"]):
        if s < d["r"].shape[0]:
            dev = np.abs(r - d["r"][s]).max()
            assert dev < 1e-9, (s, dev)
        _, kk = oracle.md_step_nose(bs, r, r1, nose, kk, hist, p)
    assert kk == meta["pred"]
    print("final: r %.2e  r1 %.2e  logs %.15g (cook %.15g)  logs1 %.15g (cook %.15g)" % (
        np.abs(r - c1["cfgs"][0]["rp"]).max(), np.abs(r1 - c1["cfgs"][1]["rp"]).max(),
        nose[0], c1["cfgs"][0]["logs"], nose[1], c1["cfgs"][1]["logs"]))
    assert np.abs(r - c1["cfgs"][0]["rp"]).max() < 1e-9
    assert np.abs(r1 - c1["cfgs"][1]["rp"]).max() < 1e-10
    assert nose[0] == pytest.approx(c1["cfgs"][0]["logs"], rel=1e-9)
    assert nose[1] == pytest.approx(c1["cfgs"][1]["logs"], rel=1e-9)


def test_predictor_matters(case):
    """a mistake planted in the oracle's input: without the predictor
    history (k = 0 throughout) the run leaves cook's trajectory"""
    d, meta, (c0, c1), _ = case
    bs = oracle.fixture_system(meta)
    r = c0["cfgs"][0]["rp"].copy()
    r1 = c0["cfgs"][1]["rp"].copy()
    nose = np.array([0.0, 0.0, 0.0, 0.0])
    p = nose_step_params(meta)
    p.trvp_k = 0
    hist = np.zeros((1, 1 + r.size))
    kk = 0
    for s in range(meta["nsteps"]):
        _, kk = oracle.md_step_nose(bs, r, r1, nose, kk, hist, p)
    assert abs(nose[0] - c1["cfgs"][0]["logs"]) > 1e-6 * abs(c1["cfgs"][0]["logs"])
