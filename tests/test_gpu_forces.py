"""GPU parity of the HIP force path (both precisions) against the golden
vectors of the reference cook and the oracle, through the C ABI."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import PUNIT, engine_provider

pytestmark = pytest.mark.gpu

CASES = ["spce1000_lattice", "spce1000_md", "brine506"]
# Ewald corrections: background energy of a charged box, slab correction,
# slab with background force and diagonal term (cook fixtures)
ECASES = ["brine505_charged", "brine506_slab", "brine505_slabbg"]
# relative RMS force error bound vs the FP64 reference: 1e-5 for the mixed
# engine, 1e-10 for the plain FP64 path
TOL = {abi.MDX_FP64: 1e-10, abi.MDX_MIXED: 1e-5}


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
@pytest.mark.parametrize("name", CASES + ECASES)
def test_forces_vs_golden(name, mode):
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    print("case", name, "mode", mode, bs.info["nk"], flush=True)
    eng = Engine(bs, mode)
    f, out = eng.forces(d["r"], measure=1)
    _, oref = oracle.forces(oracle.fixture_system(meta), d["r"])
    e = rel_rms(f, d["f"])
    print("rel force err", e, "npair", out.npair, oref.npair, flush=True)
    assert e < TOL[mode]
    assert out.npair == oref.npair
    V = meta["L"] ** 3
    Epot = out.En_pot + out.En_el
    etol = 1e-10 if mode == abi.MDX_FP64 else 2e-6
    assert abs(Epot - meta["Epot"]) < etol * abs(meta["Epot"])
    P = [out.Pvir[i] / (3 * V) * PUNIT for i in range(3)] + [out.Pvir[i] * PUNIT for i in range(3, 6)]
    ref = meta["Ptens_printed_Pa"]
    ptol = 1e-9 if mode == abi.MDX_FP64 else 1e-5
    scale = max(abs(x) for x in ref[:3])
    for a, b in zip(P[:3], ref[:3]):
        assert abs(a - b) < ptol * scale
    # off-diagonal entries are raw (not /3V) in the printed tensor
    scale_off = scale * 3 * V
    for a, b in zip(P[3:], ref[3:]):
        assert abs(a - b) < ptol * scale_off
    # measure=0 gives the same forces
    f0, out0 = eng.forces(d["r"], measure=0)
    assert rel_rms(f0, f) < (1e-14 if mode == abi.MDX_FP64 else 1e-6)
    eng.close()


@pytest.mark.parametrize("name", CASES)
def test_cook_cell_membership(name):
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, abi.MDX_MIXED)
    cid = eng.cook_cells(d["r"])
    _, _, ocid = oracle.forces(oracle.fixture_system(meta), d["r"], cellid=True)
    assert (cid == ocid).all()
    eng.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
@pytest.mark.parametrize("w,pcols", [(1, 0), (1, 1), (2, 5), (2, 3), (2, 1), (3, 7)])
def test_stencil_and_part_split(mode, w, pcols):
    """every cell stencil width and split of the stencil columns over
    parts gives the reference forces and the same pair census"""
    d, meta = oracle.load_fixture("spce1000_md")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, mode)
    f0, out0 = eng.forces(d["r"], measure=1)
    eng.set_stencil(w, pcols)
    info = eng.info()
    print("w", w, "pcols", pcols, info["stencil"], info["ncx"], flush=True)
    f, out = eng.forces(d["r"], measure=1)
    assert info["stencil"] == w and info["pair_cols"] == pcols
    assert rel_rms(f, d["f"]) < TOL[mode]
    assert out.npair == out0.npair and out.nlj == out0.nlj
    assert abs(out.En_pot + out.En_el - meta["Epot"]) < (1e-10 if mode == abi.MDX_FP64 else 2e-6) * abs(meta["Epot"])
    for a, b in zip(out.Pvir, out0.Pvir):
        assert abs(a - b) < 1e-9 * max(abs(x) for x in out0.Pvir[:3]) * (1 if mode == abi.MDX_FP64 else 1e3)
    eng.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_forces_are_deterministic(mode):
    """the cell sort orders each cell's sites by index and every sum has a
    fixed order, so two calls give the same bits; the w=1 grid puts more
    than 64 sites in a cell (the rank-sort branch)"""
    d, meta = oracle.load_fixture("spce1000_md")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, mode)
    for w in (2, 1):
        eng.set_stencil(w, 0)
        f1, o1 = eng.forces(d["r"], measure=1)
        f2, o2 = eng.forces(d["r"], measure=1)
        assert np.array_equal(f1, f2)
        assert o1.En_el == o2.En_el and o1.En_pot == o2.En_pot and list(o1.Pvir) == list(o2.Pvir)
        assert rel_rms(f1, d["f"]) < TOL[mode]
    eng.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
@pytest.mark.parametrize("name,parts", [("spce1000_md", 0), ("brine506", 3), ("polar306", 0)])
def test_all_pair_path(name, parts, mode):
    """the all-pair path (cook/pair.c, minimum image per site pair) gives
    the reference forces and the same census as the cell path"""
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, mode)
    r = d["r"] if name != "polar306" else np.concatenate([d["r"], d["rpol"][0]])
    fr = d["f"] if name != "polar306" else np.concatenate([d["f"][0], d["fpol"][0]])
    f0, o0 = eng.forces(r, measure=1)
    eng.set_stencil(0, parts)
    assert eng.info()["stencil"] == 0
    f1, o1 = eng.forces(r, measure=1)
    print(name, "all-pair rel err", rel_rms(f1, fr), "cells", rel_rms(f0, fr), flush=True)
    assert rel_rms(f1, fr) < TOL[mode]
    assert o1.npair == o0.npair and o1.nlj == o0.nlj
    assert abs((o1.En_el + o1.En_pot) - (o0.En_el + o0.En_pot)) < 1e-9 * abs(o0.En_el + o0.En_pot) * (1 if mode == abi.MDX_FP64 else 1e3)
    eng.close()


@pytest.mark.parametrize("name", ["spce1000_lattice", "brine506"])
def test_kspace_pass2_on_matrix_cores(name):
    """the mixed engine's pass 2 on the matrix cores gives cook's forces
    within the mixed bar and agrees with the FMA variant to FP32 rounding,
    for every kz chunk and part split and each choice of products
    (energies do not depend on pass 2)"""
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, abi.MDX_MIXED)
    assert eng.info()["kmfma"] == 1 and eng.info()["kprod"] == 6     # the defaults
    for kzc in (8, 16, 24, 32, 48, 64):
        for parts in (1, 5, 16):
            eng.set_kchunk(kzc, parts)
            # FP32 operands, split-FP16 operands with 3 or 6 products (the
            # default pass 1 takes its operand parts from the same setting)
            for kp in (0, 3, 6):
                eng.set_kspace_products(kp)
                eng.set_kspace_mfma(0)
                f0, o0 = eng.forces(d["r"], 1)
                for tiles in (1, 2):
                    # 1 or 2 tiles of 32 charges per wave: the same products
                    eng.set_kspace_mfma(tiles)
                    assert eng.info()["kmfma"] == tiles and eng.info()["kprod"] == kp
                    f1, o1 = eng.forces(d["r"], 1)
                    e, d01 = rel_rms(f1, d["f"]), rel_rms(f1, f0)
                    print(name, kzc, parts, kp, tiles, "mfma vs cook %.2e, vs fma %.2e" % (e, d01),
                          flush=True)
                    assert e < TOL[abi.MDX_MIXED] and d01 < 2e-6
                    assert o1.En_el == o0.En_el and o1.npair == o0.npair
                    if tiles == 1:
                        f_one = f1
                    else:
                        assert np.array_equal(f1, f_one)
    e64 = Engine(bs, abi.MDX_FP64)
    with pytest.raises(RuntimeError):
        e64.set_kspace_mfma(True)
    with pytest.raises(RuntimeError):
        e64.set_kspace_products(6)
    with pytest.raises(RuntimeError):
        eng.set_kspace_products(4)
    e64.close()
    eng.close()
    e64.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_engine_from_cook_input_files(mode, tmp_path):
    """the engine built from the .ble force field and .def deck the brine
    fixture came from (read by macsimus_amd.cookio) gives cook's forces"""
    import test_cookio
    from macsimus_amd import cookio
    d, meta = oracle.load_fixture("brine506")
    species, types, _ = test_cookio.CASES["brine506"]()
    ble, dfn = test_cookio.write_deck(str(tmp_path), meta, species, types, None)
    bs, _ = cookio.system_from_cook(ble, dfn, provider=engine_provider())
    eng = Engine(bs, mode)
    f, out = eng.forces(d["r"], 1)
    e = rel_rms(f, d["f"])
    assert e < TOL[mode], e
    assert abs(out.En_pot + out.En_el - meta["Epot"]) < (1e-10 if mode == abi.MDX_FP64 else 2e-6) * abs(meta["Epot"])
    eng.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_engine_with_nbfix(mode, tmp_path):
    """the brine deck with an Na-O nbfix (cook/simdef.c:772-798), read by
    cookio: the engine gives cook's own forces and Epot of that run
    (tests/golden/brine506_nbfix), the oracle's census, and not the forces
    of the deck without the nbfix"""
    import test_cookio
    from macsimus_amd import cookio
    d, meta = oracle.load_fixture("brine506_nbfix")
    species, types, _ = test_cookio.CASES["brine506"]()
    ble, dfn = test_cookio.write_deck(str(tmp_path), meta, species, types, None,
                                      [tuple(x) for x in meta["nbfix"]])
    print("brine506 + NA-O nbfix mode %d" % mode, flush=True)
    bs, _ = cookio.system_from_cook(ble, dfn, provider=engine_provider())
    bso, _ = cookio.system_from_cook(ble, dfn, provider=oracle.provider())
    f_ref, o_ref = oracle.forces(bso, d["r"], 1)
    eng = Engine(bs, mode)
    f, out = eng.forces(d["r"], 1)
    eng.close()
    assert rel_rms(f, d["f"]) < TOL[mode]
    assert abs(out.En_pot + out.En_el - meta["Epot"]) < (1e-10 if mode == abi.MDX_FP64 else 2e-6) * abs(meta["Epot"])
    assert out.nlj == o_ref.nlj and out.npair == o_ref.npair
    d0, _ = oracle.load_fixture("brine506")
    assert rel_rms(d0["f"], d["f"]) > 1e-4       # another run: the nbfix changes it


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_resize_to_another_system_and_back(mode):
    """mdx_resize: the context takes another system (other numbers of
    sites, types and box) and gives a fresh context's bits for it with the
    same settings; back to the first system, its bits again"""
    da, ma = oracle.load_fixture("spce1000_md")
    db, mb = oracle.load_fixture("brine506")
    bsa = oracle.fixture_system(ma, provider_=engine_provider())
    bsb = oracle.fixture_system(mb, provider_=engine_provider())
    e = Engine(bsa, mode)
    e.set_stencil(2, 5)
    e.set_kchunk(16, 8)
    fa, oa = e.forces(da["r"], 1)
    e.resize(bsb)
    info = e.info()
    assert info["pair_cols"] == 5 and info["kchunk"] == 16
    fb, ob = e.forces(db["r"], 1)
    fresh = Engine(bsb, mode)
    fresh.set_stencil(2, 5)
    fresh.set_kchunk(16, 8)
    fb2, ob2 = fresh.forces(db["r"], 1)
    fresh.close()
    assert np.array_equal(fb, fb2) and ob.En_el == ob2.En_el and ob.npair == ob2.npair
    e.resize(bsa)
    fa2, oa2 = e.forces(da["r"], 1)
    e.close()
    assert np.array_equal(fa, fa2) and oa.En_pot == oa2.En_pot
