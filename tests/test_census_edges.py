"""Edge cases of cook's cell rule and pair census on the CPU: the
configurations of census_edges.py are at the edges they are meant to be
at (the oracle's census there is what the GPU tests hold the engine to)."""
import numpy as np
import pytest

import census_edges as ce
import oracle


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sharp_pairs_sit_on_the_cutoff_across_the_boundary(axis):
    ps = ce.sharp_pairs(4, axis=axis)
    assert len(ps) == 4
    bs = ce.ion_system(1, 1)
    for r in ps:
        # across the periodic boundary along axis, at the cutoff to ~1e-15
        d = r[0] - r[1]
        d -= ce.L * np.rint(d / ce.L)
        assert abs(np.sqrt((d ** 2).sum()) - ce.CUTOFF) < 1e-13
        assert abs(r[0, axis] - r[1, axis]) > ce.L / 2
        # cook's arithmetic and the plain minimum image disagree here
        _, o = oracle.forces(bs, r, measure=0, nthreads=1)
        assert o.npair != ce.naive_npair(r)


def test_edge_sites_cells():
    """cook's cells of coordinates at 0, -1e-12, L-ulp, L, k L/nc and
    beyond L: in range, the site inside its cell up to the shaved rL's few
    ulps, negative coordinates in cell 0 (not wrapped), coordinates >= L
    wrapped by subtraction"""
    bs = ce.ion_system(6, 6)
    r = ce.edge_sites()
    _, o, cid = oracle.forces(bs, r, cellid=True)
    nc = np.array(ce.NCELL)
    ic = np.stack([cid // (nc[1] * nc[2]), (cid // nc[2]) % nc[1], cid % nc[2]], 1)
    assert ((ic >= 0) & (ic < nc)).all()
    x = r[:, 0].copy()
    while (x >= ce.L).any():
        x = np.where(x >= ce.L, x - ce.L, x)
    cw = ce.L / nc[0]
    assert ic[1, 0] == 0 and x[1] < 0                       # -1e-12 stays negative
    assert ic[2, 0] == nc[0] - 1 and ic[3, 0] == 0          # L - ulp, L
    assert ic[10, 0] == nc[0] - 1                            # 2L - ulp
    for i in range(len(x)):
        assert ic[i, 0] * cw - 1e-12 <= x[i] <= (ic[i, 0] + 1) * cw + 1e-12
    # the census of these sites is the brute-force minimum-image count
    w = r - ce.L * np.floor(r / ce.L)
    n = 0
    for i in range(len(w)):
        d = w[i] - w[i + 1:]
        d -= ce.L * np.rint(d / ce.L)
        n += int(((d ** 2).sum(1) < ce.CUTOFF ** 2).sum())
    assert n == o.npair
