"""The plain Berendsen thermostat of the device MD step (sim/thermo.c:
141-160) against cook's own run (tests/golden/spce216_ber): from the
configuration cook saved (cfg0), 8 steps with thermostat="Berendsen",
tau.T = 0.05 ps end on the configuration cook saved after them; the
measured steps' Berendsen factor equals the oracle's, and the captured
graph gives the eager steps' bits."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider
from test_oracle_md import berendsen_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-10), (abi.MDX_MIXED, 2e-5)])
def test_engine_ends_on_cooks_configuration(tmp_path, mode, tol):
    meta, cfgs, p = berendsen_case(tmp_path)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    c0, c1 = cfgs["cfg0"], cfgs["cfg_ber"]
    e = Engine(bs, mode)
    e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"])
    lam = []
    for _ in range(meta["nsteps"]):
        out = e.md_steps(p, 1)
        lam.append(out.lambda_)
    r, r1 = e.download_state()
    e.close()
    dr = np.abs(r - c1["cfgs"][0]["rp"]).max()
    dr1 = np.abs(r1 - c1["cfgs"][1]["rp"]).max()
    print("Berendsen mode %d: r %.2e r1 %.2e" % (mode, dr, dr1), flush=True)
    assert dr < tol and dr1 < tol
    # the factor of each step is the oracle's
    bso = oracle.fixture_system(meta)
    ro, r1o = c0["cfgs"][0]["rp"].copy(), c0["cfgs"][1]["rp"].copy()
    for k in range(meta["nsteps"]):
        oo = oracle.md_step(bso, ro, r1o, p, measure=1)
        assert abs(lam[k] - oo.lambda_) < (1e-12 if mode == abi.MDX_FP64 else 1e-6)


def test_berendsen_graph_gives_eager_bits(tmp_path):
    meta, cfgs, p = berendsen_case(tmp_path)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    c0 = cfgs["cfg0"]
    p.measure_every = 4
    res = []
    for graph in (True, False):
        e = Engine(bs, abi.MDX_MIXED)
        e.set_graph(graph)
        e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"])
        e.md_steps(p, 12)
        res.append(e.download_state())
        e.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
