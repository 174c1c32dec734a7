"""Nose-Hoover with the time-reversible velocity predictor on the GPU
against cook's own run (tests/golden/spce216_nose): from the configuration
cook saved (cfg0), the engine's SHAKE + Nose steps follow the dumped force
calls and end on cook's final configuration and Nose variables (cfg1);
the captured step (graph, once the predictor has its full length) gives
the eager steps' bits; the mixed engine stays within its error bound."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, cookio, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    d, meta = oracle.load_fixture("spce216_nose")
    td = tmp_path_factory.mktemp("nose")
    cfgs = []
    for nm in ("cfg0", "cfg1"):
        p = td / (nm + ".cfg")
        p.write_bytes(d[nm].tobytes())
        cfgs.append(cookio.read_cfg(p))
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    nf = workloads.degrees_of_freedom(meta["forcefield"], meta["counts"])
    p = abi.Step(h=meta["h"], eps=1e-12, maxit=0, thermostat=2, T=meta["T"], tau=meta["tauT"],
                 nf=float(nf), measure_every=1, trvp_k=meta["pred"])
    return d, meta, cfgs, bs, p


def start(bs, c0, mode, graph=True):
    e = Engine(bs, mode)
    e.set_graph(graph)
    e.upload_state(c0["cfgs"][0]["rp"], c0["cfgs"][1]["rp"])
    e.set_nose(c0["cfgs"][0]["logs"], c0["cfgs"][1]["logs"], 0, 0)
    return e


@pytest.mark.parametrize("mode,tol", [(abi.MDX_FP64, 1e-9), (abi.MDX_MIXED, 2e-5)])
def test_engine_follows_cooks_nose_run(case, mode, tol):
    d, meta, (c0, c1), bs, p = case
    print("spce216_nose mode %d" % mode, flush=True)
    e = start(bs, c0, mode)
    for s in range(meta["nsteps"]):
        r, _ = e.download_state()
        if s < d["r"].shape[0]:
            assert np.abs(r - d["r"][s]).max() < tol, s
        so = e.md_steps(p, 1)
    r, r1 = e.download_state()
    logs, logs1, k, kk, _ = e.get_nose()
    e.close()
    print("  final r %.2e r1 %.2e logs %.15g (cook %.15g) logs1 %.15g (cook %.15g)" % (
        np.abs(r - c1["cfgs"][0]["rp"]).max(), np.abs(r1 - c1["cfgs"][1]["rp"]).max(),
        logs, c1["cfgs"][0]["logs"], logs1, c1["cfgs"][1]["logs"]), flush=True)
    assert (k, kk) == (meta["pred"], meta["pred"])
    assert np.abs(r - c1["cfgs"][0]["rp"]).max() < tol
    assert logs == pytest.approx(c1["cfgs"][0]["logs"], rel=100 * tol)
    assert logs1 == pytest.approx(c1["cfgs"][1]["logs"], rel=100 * tol)
    assert so.logs == pytest.approx(logs, rel=1e-15)
    assert np.isfinite(so.En_ext) and so.En_ext != 0


def test_graph_replay_gives_eager_bits(case):
    """the same 8 steps in one call (steps after the predictor has grown
    replay a captured graph) and step by step eagerly, then 8 more"""
    d, meta, (c0, c1), bs, p = case
    p2 = abi.Step.from_buffer_copy(p)
    p2.measure_every = 4
    ea = start(bs, c0, abi.MDX_FP64, graph=True)
    eb = start(bs, c0, abi.MDX_FP64, graph=False)
    for rnd in range(2):
        oa = ea.md_steps(p2, 8)
        ob = eb.md_steps(p2, 8)
        (ra, r1a), (rb, r1b) = ea.download_state(), eb.download_state()
        assert np.array_equal(ra, rb) and np.array_equal(r1a, r1b), rnd
        na, nb = ea.get_nose(), eb.get_nose()
        assert na[0] == nb[0] and na[1] == nb[1] and np.array_equal(na[4], nb[4])
        assert oa.En_ext == ob.En_ext
    ea.close()
    eb.close()


def test_nose_state_round_trip_and_errors(case):
    d, meta, (c0, c1), bs, p = case
    e = start(bs, c0, abi.MDX_FP64)
    e.md_steps(p, 2)
    logs, logs1, k, kk, hist = e.get_nose()
    assert (k, kk) == (3, 2) and np.abs(hist[:2]).max() > 0 and not hist[2].any()
    e.set_nose(logs, logs1, k, kk, hist)
    assert np.array_equal(e.get_nose()[4], hist)
    bad = abi.Step.from_buffer_copy(p)
    bad.trvp_k = 12
    with pytest.raises(RuntimeError):
        e.md_steps(bad, 1)
    e.close()
