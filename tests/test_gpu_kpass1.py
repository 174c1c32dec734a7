"""The k-space pass-1 variants of the mixed mode (FP32 FMA chains, the FP32
matrix cores with 16 or 8 tiles of 4 k rows per wave or as row units, and
the FP16 matrix cores with split operands in four unit shapes, each with
3 or 6 products) give the same structure-factor sums: forces of cook's
brine506 positions against the oracle, and the 194,304-site brine (93 k
segments: a partial workgroup for 8 tiles) of each variant against the FMA
variant and FP64."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import build_system, engine_provider

pytestmark = pytest.mark.gpu

TILES = (0, 1, 2, 16, 8, 4, 5, 6, 7)


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_pass1_variants_vs_oracle():
    d, meta = oracle.load_fixture("brine506")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    f_ref, o_ref = oracle.forces(oracle.fixture_system(meta), d["r"], 1)
    e = Engine(bs, abi.MDX_MIXED)
    assert e.info()["kpass1_tiles"] == 5          # the default without Drudes
    for kp in (6, 3):
        e.set_kspace_products(kp)
        for t in TILES:
            print("brine506 pass-1 tiles %d products %d" % (t, kp), flush=True)
            e.set_kspace_pass1(t)
            assert e.info()["kpass1_tiles"] == t
            f, out = e.forces(d["r"], 1)
            assert rel_rms(f, f_ref) < 1e-5
            assert abs(out.En_el - o_ref.En_el) < 2e-6 * abs(o_ref.En_el + o_ref.En_pot)
    for bad in (3, 9, -1):
        with pytest.raises(RuntimeError):
            e.set_kspace_pass1(bad)
    e.close()
    e = Engine(bs, abi.MDX_FP64)
    with pytest.raises(RuntimeError):
        e.set_kspace_pass1(16)                  # MDX_MIXED only
    e.close()


def test_pass1_variants_full_size():
    wl = workloads.brine_c2()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    e = Engine(bs, abi.MDX_FP64)
    f64, o64 = e.forces(wl["r"], 1)
    e.close()
    e = Engine(bs, abi.MDX_MIXED)
    res = {}
    for t in TILES:
        print("C2 pass-1 tiles %d" % t, flush=True)
        e.set_kspace_pass1(t)
        res[t] = e.forces(wl["r"], 1)
    # the split-FP16 units share their operands' bits across shapes: the
    # same sums in each shape up to the order of the chunk sums
    e.set_kspace_products(3)
    for t in (4, 5):
        print("C2 pass-1 tiles %d, 3 products" % t, flush=True)
        e.set_kspace_pass1(t)
        res[(t, 3)] = e.forces(wl["r"], 1)
    nseg = e.info()["ksegments"]
    e.close()
    print("k segments %d" % nseg, flush=True)
    for t in list(TILES) + [(4, 3), (5, 3)]:
        f, out = res[t]
        err = rel_rms(f, f64)
        print("  tiles %s: rel RMS force err vs FP64 %.3e" % (t, err), flush=True)
        assert err < 1e-5
        assert rel_rms(f, res[0][0]) < 1e-6
        E64 = o64.En_pot + o64.En_el
        assert abs(out.En_pot + out.En_el - E64) < 1e-6 * abs(E64)
        assert abs(out.En_el - res[0][1].En_el) < 1e-6 * abs(E64)
