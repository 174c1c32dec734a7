"""RDF histograms of cook's measure path (cook/ljqqnp.c:218-225) in the
oracle, pinned by cook's own binary histograms.

tests/golden/brine506_rdf holds a 3-step MD run of 3 NaCl + 500 SPC/E
with rdf.grid=20, rdf.cutoff=9 (the r-space cutoff): every force call
dumped and the .rdf records cook saved at the end (nmeas=2: the initial
evaluation, dump 0, and the measured last step of the cycle, dump 3;
dump 1 repeats dump 0's positions).  The oracle's histograms of those two
calls must equal cook's bin for bin.  Each dump alone is pinned by a brute
force count over all pairs with the minimum image (no cells, no cook
code), which the fixture's sum alone does not fix.
"""
import numpy as np
import pytest

import oracle

MEASURED = (0, 3)


@pytest.fixture(scope="module")
def case():
    d, meta = oracle.load_fixture("brine506_rdf")
    bs = oracle.fixture_system(meta)
    return d, meta, bs


def brute_hist(bs, r, L, grid, cutoff, merged=False):
    """all pairs i<j, minimum image, same-molecule EXCL pairs skipped,
    rr < cutoff^2 of the pair loop, bin int(r*grid) below nhist"""
    ns = bs.struct.ns
    types = np.array(bs.info["site_type"])
    mol = np.repeat(np.arange(len(bs.info["mol_first"]) - 1), np.diff(bs.info["mol_first"]))
    first = np.array(bs.info["mol_first"])[mol]
    excl = np.array(bs.info["excl"], np.int64)
    nt = bs.struct.ntypes
    nrec, nh = oracle.rdf_shape(nt, -grid if merged else grid, cutoff)
    hist = np.zeros((nrec, nh), np.uint64)
    cq = bs.struct.cutoff ** 2
    for i in range(ns - 1):
        d = r[i] - r[i + 1:]
        d -= L * np.rint(d / L)
        rr = (d * d).sum(1)
        j = np.arange(i + 1, ns)
        same = mol[j] == mol[i]
        ex = same & (((excl[i] >> (j - first[i])) & 1) == 1)
        ok = (rr < cq) & ~ex
        b = (np.sqrt(rr[ok]) * abs(grid)).astype(np.int64)
        ti, tj = np.maximum(types[i], types[j[ok]]), np.minimum(types[i], types[j[ok]])
        p = np.zeros_like(ti) if merged else ti * (ti + 1) // 2 + tj
        keep = b < nh
        np.add.at(hist, (p[keep], b[keep]), 1)
    return hist


def cook_records(d, meta):
    nt = meta["rdf_nsites"]
    out = np.zeros((nt * (nt + 1) // 2, meta["rdf_pairs"][0]["nhist"]), np.uint64)
    for i in range(nt):
        for j in range(i + 1):
            out[i * (i + 1) // 2 + j] = d["hist_%d_%d" % (i, j)]
    return out


def test_rdf_records_equal_cooks(case):
    d, meta, bs = case
    grid, cut = meta["rdf_grid"], meta["rdf_cutoff"]
    for p in meta["rdf_pairs"]:
        assert p["nmeas"] == len(MEASURED)
        assert p["V"] == pytest.approx(len(MEASURED) * meta["L"] ** 3, rel=1e-12)
    hist = None
    for k in MEASURED:
        _, _, hist = oracle.forces_rdf(bs, d["r"][k], grid, cut, hist)
    ref = cook_records(d, meta)
    assert hist.shape == ref.shape
    assert int(ref.sum()) > 400000
    np.testing.assert_array_equal(hist, ref)


def test_rdf_each_call_equals_brute_force(case):
    d, meta, bs = case
    grid, cut = meta["rdf_grid"], meta["rdf_cutoff"]
    L = np.array([meta["L"]] * 3)
    for k in MEASURED:
        _, _, h = oracle.forces_rdf(bs, d["r"][k], grid, cut)
        np.testing.assert_array_equal(h, brute_hist(bs, d["r"][k], L, grid, cut))


def test_rdf_merged_and_short_records(case):
    """grid<0: one record, the sum of the type-pair records; a shorter
    rdf.cutoff keeps the leading bins; the grid is truncated to an integer
    (abs() of a double in initrdf)"""
    d, meta, bs = case
    r = d["r"][0]
    _, _, full = oracle.forces_rdf(bs, r, 20, 9.0)
    _, _, merged = oracle.forces_rdf(bs, r, -20, 9.0)
    np.testing.assert_array_equal(merged[0], full.sum(0))
    _, _, short = oracle.forces_rdf(bs, r, 20, 4.0)
    assert short.shape == (full.shape[0], 81)
    np.testing.assert_array_equal(short, full[:, :81])
    assert oracle.rdf_shape(4, 20.7, 9.0) == oracle.rdf_shape(4, 20, 9.0)
    with pytest.raises(RuntimeError):
        oracle.forces_rdf(bs, r, 0, 9.0)


def test_rdf_file_writer_gives_cooks_bytes(case, tmp_path):
    """write_rdf of the oracle's histograms of the measured calls, with
    cook's nmeas and V, is byte for byte the .rdf file cook saved (npair of
    every record from cutcorr()'s count, names, site counts, checksums)"""
    from macsimus_amd import cookio
    d, meta, bs = case
    grid, cut = meta["rdf_grid"], meta["rdf_cutoff"]
    hist = None
    for k in MEASURED:
        _, _, hist = oracle.forces_rdf(bs, d["r"][k], grid, cut, hist)
    p0 = meta["rdf_pairs"][0]
    path = tmp_path / "sys.rdf"
    cookio.write_rdf(path, dict(hist=hist, nmeas=p0["nmeas"], V=p0["V"], grid=grid),
                     meta["forcefield"], meta["counts"])
    assert path.read_bytes() == d["rdf_file"].tobytes()
    g, nsites, recs = cookio.read_rdf(path)
    assert (g, nsites, len(recs)) == (grid, meta["rdf_nsites"], len(meta["rdf_pairs"]))
    for r, p in zip(recs, meta["rdf_pairs"]):
        for key in ("names", "indx", "ns", "npair", "grid", "V", "nmeas"):
            assert r[key] == p[key], key
        np.testing.assert_array_equal(r["hist"], d["hist_%d_%d" % tuple(p["indx"])])


def test_rdf_file_rejects_damage_and_overflow(case, tmp_path):
    from macsimus_amd import cookio
    d, meta, _ = case
    raw = bytearray(d["rdf_file"].tobytes())
    raw[200] ^= 1
    bad = tmp_path / "bad.rdf"
    bad.write_bytes(bytes(raw))
    with pytest.raises(cookio.CookInputError):
        cookio.read_rdf(bad)
    h = np.zeros((10, 181), np.uint64)
    h[0, 0] = 1 << 32
    with pytest.raises(ValueError):
        cookio.write_rdf(tmp_path / "x.rdf", dict(hist=h, nmeas=1, V=1.0, grid=20),
                         meta["forcefield"], meta["counts"])
    # merged record: one record of all pairs, npair the total
    np1 = cookio.rdf_npair(meta["forcefield"], meta["counts"], merged=True)
    assert np1[0] == sum(p["npair"] for p in meta["rdf_pairs"])
