"""The engine's pair census and cook cells equal cook's (the oracle's
restatement of cook/lc.c) at the edges: pairs at the cutoff to an ulp
across each periodic boundary, and sites on cell faces, at 0, at -1e-12,
at L - ulp, at L and beyond.  The engine decides pairs within rounding of
the cutoff with cook's own arithmetic, so these agree by construction."""
import numpy as np
import pytest

import census_edges as ce
import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_census_of_sharp_pairs(axis, mode):
    ps = ce.sharp_pairs(6, axis=axis)
    bso = ce.ion_system(1, 1)
    eng = Engine(ce.ion_system(1, 1, engine_provider()), mode)
    print("axis", axis, "mode", mode, len(ps), flush=True)
    for r in ps:
        _, oo = oracle.forces(bso, r, measure=1, nthreads=1)
        _, oe = eng.forces(r, measure=1)
        _, oe0 = eng.forces(r, measure=0)
        assert oe.npair == oo.npair and oe0.npair == oo.npair, r.tolist()
    eng.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_rdf_of_sharp_pairs(mode):
    """the RDF records count the pair as cook does (in its last bin when
    cook has it inside the cutoff)"""
    grid = 20
    bso = ce.ion_system(1, 1)
    eng = Engine(ce.ion_system(1, 1, engine_provider()), mode)
    eng.set_rdf(grid, ce.CUTOFF)
    hist = None
    for axis in (0, 1, 2):
        for r in ce.sharp_pairs(4, axis=axis, seed=11):
            _, _, hist = oracle.forces_rdf(bso, r, grid, ce.CUTOFF, hist=hist, nthreads=1)
            eng.forces(r, measure=1)
    h = eng.rdf()["hist"]
    assert h.sum() > 0
    assert np.array_equal(h.reshape(hist.shape), hist)
    eng.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_cells_and_census_of_edge_sites(mode):
    bso = ce.ion_system(6, 6)
    eng = Engine(ce.ion_system(6, 6, engine_provider()), mode)
    r = ce.edge_sites()
    _, oo, cid = oracle.forces(bso, r, cellid=True)
    assert np.array_equal(eng.cook_cells(r), cid)
    _, oe = eng.forces(r, measure=1)
    assert oe.npair == oo.npair
    eng.close()
