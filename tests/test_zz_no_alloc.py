"""mdx_adjoint_lj_noalloc (scratch in a static device array, no device
allocation) against mdx_adjoint_lj (scratch allocated per call) through
the C ABI: the same kernels in the same order, so the same bits, on two
calls in a row with a smaller system on the second; and a system beyond
the array is refused with MDX_E_ARG."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.engine import Engine
from macsimus_amd.system import build_system, engine_provider

pytestmark = pytest.mark.gpu


def _check(name, mode, seed):
    d, meta = oracle.load_fixture(name)
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    eng = Engine(bs, mode)
    r = d["r"] if d["r"].ndim == 2 else d["r"][0]
    w = np.random.default_rng(seed).normal(size=(bs.struct.ns, 3))
    print("case", name, "mode", mode, "ns", bs.struct.ns, flush=True)
    gF1, gE1 = eng.adjoint_lj(r, w)
    gF2, gE2 = eng.adjoint_lj_noalloc(r, w)
    eng.close()
    assert np.abs(gF1).max() > 0 and np.abs(gE1).max() > 0
    assert np.array_equal(gF1, gF2) and np.array_equal(gE1, gE2)
    return bs.struct.ns


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_noalloc_adjoint_equals_allocating_one(mode):
    n1 = _check("brine506", mode, 1)
    n2 = _check("spce216_nve", mode, 2)        # second call: a smaller system
    assert n2 < n1
    n3 = _check("brine506", mode, 3)        # and back to the larger one
    assert n3 == n1


def test_noalloc_adjoint_refuses_a_system_beyond_its_scratch():
    _, meta = oracle.load_fixture("brine506")
    # 12,000 waters: 36,000 sites x (3 + 8 x 4 types) doubles > 8 MiB
    bs = build_system(meta["forcefield"], counts=[0, 0, 12000], L=[71.2] * 3, cutoff=9.0,
                      provider=engine_provider())
    eng = Engine(bs, abi.MDX_MIXED)
    r = np.zeros((bs.struct.ns, 3))
    with pytest.raises(RuntimeError, match="MDX_ADJ_SCRATCH_BYTES"):
        eng.adjoint_lj_noalloc(r, r)
    eng.close()
