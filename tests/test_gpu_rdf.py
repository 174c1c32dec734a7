"""RDF histograms of the engine's measured calls (cook/ljqqnp.c:218-225)
against cook's own records and the oracle, bin for bin.

- brine506_rdf: cook's histograms of its two measured force calls (dumps
  0 and 3) are the engine's, in both modes, on the cell and all-pair
  paths; unmeasured calls add nothing; nmeas and V as cook's.
- merged (grid < 0) and shorter rdf.cutoff forms, a measured MD step and a
  Drude system against the oracle.
- 194,304 sites: the records hold every pair of the pair census that is
  not excluded (each SPC/E molecule has 3 excluded pairs), the same in
  both modes.
"""
import time

import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import build_system, engine_provider

pytestmark = pytest.mark.gpu

MODES = [abi.MDX_FP64, abi.MDX_MIXED]
MEASURED = (0, 3)


@pytest.fixture(scope="module")
def brine():
    d, meta = oracle.load_fixture("brine506_rdf")
    nt = meta["rdf_nsites"]
    ref = np.zeros((nt * (nt + 1) // 2, meta["rdf_pairs"][0]["nhist"]), np.uint64)
    for i in range(nt):
        for j in range(i + 1):
            ref[i * (i + 1) // 2 + j] = d["hist_%d_%d" % (i, j)]
    return d, meta, ref, oracle.fixture_system(meta, provider_=engine_provider())


@pytest.mark.parametrize("w", [2, 0])
@pytest.mark.parametrize("mode", MODES)
def test_rdf_equals_cooks_records(brine, mode, w):
    d, meta, ref, bs = brine
    print("brine506_rdf mode %d stencil %d" % (mode, w), flush=True)
    e = Engine(bs, mode)
    e.set_stencil(w)
    e.set_rdf(meta["rdf_grid"], meta["rdf_cutoff"])
    e.forces(d["r"][MEASURED[0]], 1)
    e.forces(d["r"][2], 0)                      # unmeasured: not counted
    e.forces(d["r"][MEASURED[1]], 1)
    h = e.rdf()
    e.close()
    np.testing.assert_array_equal(h["hist"], ref)
    p0 = meta["rdf_pairs"][0]
    assert h["nmeas"] == p0["nmeas"] == 2
    assert h["V"] == pytest.approx(p0["V"], rel=1e-15)


@pytest.mark.parametrize("mode", MODES)
def test_rdf_forms_and_md_step(brine, mode):
    d, meta, ref, bs = brine
    bso = oracle.fixture_system(meta)
    e = Engine(bs, mode)
    # records x bins: 10 x 1594 = 15,940 (LDS histograms at nearly their
    # 64 KiB budget) and 10 x 18,001 (past it: global 64-bit atomics)
    for grid, cut in ((-20, 9.0), (20, 4.0), (33, 12.0), (177, 9.0), (2000, 9.0)):
        print("rdf grid %d cutoff %g" % (grid, cut), flush=True)
        e.set_rdf(grid, cut)
        e.forces(d["r"][3], 1)
        _, _, want = oracle.forces_rdf(bso, d["r"][3], grid, cut)
        got = e.rdf()
        assert got["hist"].shape == want.shape and got["nmeas"] == 1
        np.testing.assert_array_equal(got["hist"], want)
    # a measured MD step counts the pairs of its force call, at the
    # positions the step starts from; set_rdf restarts from zero
    e.set_rdf(20, 9.0)
    e.upload_state(d["r"][0], np.zeros_like(d["r"][0]))
    p = abi.Step(h=meta["h"], eps=1e-10, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=1.0,
                 measure_every=0)
    e.md_steps(p, 1)
    r_next = e.download_state()[0]
    e.md_steps(p, 1)
    got = e.rdf()
    e.close()
    L = meta["L"]
    _, _, want = oracle.forces_rdf(bso, d["r"][0], 20, 9.0)
    _, _, want = oracle.forces_rdf(bso, np.mod(r_next, L), 20, 9.0, want)
    assert got["nmeas"] == 2
    np.testing.assert_array_equal(got["hist"], want)


@pytest.mark.parametrize("mode", MODES)
def test_rdf_of_drude_system(mode):
    """for Drude sites the centres are counted (cook/ljqqpol.c:477-485)"""
    d, meta = oracle.load_fixture("polar306")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    bso = oracle.fixture_system(meta)
    r = np.concatenate([d["r"], d["rpol"][0]])
    e = Engine(bs, mode)
    e.set_rdf(25, meta["cutoff"])
    e.forces(r, 1)
    got = e.rdf()
    e.close()
    _, _, want = oracle.forces_rdf(bso, r, 25, meta["cutoff"])
    assert want.sum() > 0
    np.testing.assert_array_equal(got["hist"], want)


def test_rdf_bad_arguments(brine):
    d, meta, ref, bs = brine
    e = Engine(bs, abi.MDX_FP64)
    with pytest.raises(RuntimeError):
        e.rdf()                                  # off
    with pytest.raises(RuntimeError):
        e.set_rdf(20, -1.0)
    with pytest.raises(RuntimeError):
        e.set_rdf(20000, 9.0)                    # > 100000 bins
    e.set_rdf(20, 9.0)
    e.set_rdf(0, 0.0)                            # off again
    with pytest.raises(RuntimeError):
        e.rdf()
    e.close()


def test_rdf_full_size_census():
    wl = workloads.brine_c2()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    nexcl = 3 * wl["counts"][2]
    res = {}
    for mode in MODES:
        print("C2 rdf mode %d" % mode, flush=True)
        e = Engine(bs, mode)
        e.forces(wl["r"], 1)
        t0 = time.perf_counter()
        e.forces(wl["r"], 1)
        t1 = time.perf_counter()
        e.set_rdf(20, wl["cutoff"])
        e.forces(wl["r"], 1)
        t2 = time.perf_counter()
        _, out = e.forces(wl["r"], 1)
        t3 = time.perf_counter()
        h = e.rdf()
        e.close()
        print("  measured call %.2f ms, with RDF %.2f ms" % ((t1 - t0) * 1e3, (t3 - t2) * 1e3),
              flush=True)
        assert h["nmeas"] == 2
        assert int(h["hist"].sum()) == 2 * (out.npair - nexcl)
        # the last bin (r = cutoff) is empty: pairs are counted below the cutoff
        assert not h["hist"][:, -1].any()
        res[mode] = h["hist"]
    np.testing.assert_array_equal(res[abi.MDX_FP64], res[abi.MDX_MIXED])
