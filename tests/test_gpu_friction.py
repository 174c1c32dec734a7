"""The decoupled friction thermostats (sim/thermo.c:107-135) in the
device MD step against cook's own runs (tests/golden/spce216_friction):
from the configuration cook saved (cfg0), 6 steps with friction on the
molecular translations (11), the intramolecular motion (12) or both (13)
end on the configuration cook saved after each; the graph-captured steps
give the eager steps' bits."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, cookio
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider
from test_oracle_friction import THERMOSTATS, step_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    d, meta = oracle.load_fixture("spce216_friction")
    td = tmp_path_factory.mktemp("fricg")
    cfgs = {}
    for nm in ("cfg0", "cfg_tr", "cfg_in", "cfg_fr"):
        p = td / (nm + ".cfg")
        p.write_bytes(d[nm].tobytes())
        cfgs[nm] = cookio.read_cfg(p)
    return d, meta, cfgs, oracle.fixture_system(meta, provider_=engine_provider())


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-10), (abi.MDX_MIXED, 2e-5)])
@pytest.mark.parametrize("tag", ["tr", "in", "fr"])
def test_engine_ends_on_cooks_configuration(case, tag, mode, tol):
    d, meta, cfgs, bs = case
    c0, c1 = cfgs["cfg0"], cfgs["cfg_" + tag]
    e = Engine(bs, mode)
    e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"])
    out = e.md_steps(step_params(meta, THERMOSTATS[tag]), meta["nsteps"])
    r, r1 = e.download_state()
    e.close()
    dr = np.abs(r - c1["cfgs"][0]["rp"]).max()
    dr1 = np.abs(r1 - c1["cfgs"][1]["rp"]).max()
    print("%s mode %d: r %.2e r1 %.2e" % (tag, mode, dr, dr1), flush=True)
    assert dr < tol and dr1 < tol
    assert 0 < out.En_kin_tr < out.En_kin


def test_friction_graph_gives_eager_bits(case):
    d, meta, cfgs, bs = case
    c0 = cfgs["cfg0"]
    p = step_params(meta, 13)
    p.measure_every = 5
    res = []
    for graph in (True, False):
        e = Engine(bs, abi.MDX_FP64)
        e.set_graph(graph)
        e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"])
        e.md_steps(p, 10)
        res.append(e.download_state())
        e.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
