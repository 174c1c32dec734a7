"""The POLAR&1 shell-core repulsion of the oracle (cook/ljqqpol.c:41-46,
525-535, 561-648), pinned without a cook run (cook builds it only for the
Buckingham and Busing forms, whose build needs a generated simopt.h):

- kappa = 0 gives the forces of the system without shell data, bit for bit;
- with every Drude on its centre the shell terms of a pair cancel
  (+kappa urep(r_cc) between the centres, -kappa urep(r_dc) with r_dc =
  r_cc), so the total force on each site (centre + Drude), the energy and
  the virial equal the kappa = 0 ones while the split between centre and
  Drude does not;
- the forces are minus the gradient of the energy: central differences of
  En_pot + En_el in a Drude offset (force on the Drude) and in a centre
  position with its Drude offset held (centre + Drude);
- a shell needs Drudes and an exponential repulsion."""
import numpy as np
import pytest

import oracle
import shellcore
from macsimus_amd.system import build_system


def energy(bs, r):
    _, out = oracle.forces(bs, r)
    return out.En_pot + out.En_el


def test_kappa_zero_is_no_shell():
    r, bs0, meta = shellcore.system(kappa=0.0)
    f0, o0 = oracle.forces(bs0, r)
    bsn = oracle.fixture_system(dict(meta, forcefield=shellcore.buck_forcefield(meta["forcefield"])))
    fn, on = oracle.forces(bsn, r)
    np.testing.assert_array_equal(f0, fn)
    assert o0.En_pot == on.En_pot and o0.En_el == on.En_el


def test_drudes_on_centres_cancel():
    r, bs, _ = shellcore.system()
    _, bs0, _ = shellcore.system(kappa=0.0)
    ns = bs.struct.ns
    r = r.copy()
    r[ns:] = 0
    f, o = oracle.forces(bs, r)
    f0, o0 = oracle.forces(bs0, r)
    tot, tot0 = f[:ns] + f[ns:], f0[:ns] + f0[ns:]
    scale = np.abs(tot0).max()
    print("site-total force diff %.2e, centre diff %.2e of %.2e" % (
        np.abs(tot - tot0).max(), np.abs(f[:ns] - f0[:ns]).max(), scale))
    assert np.abs(tot - tot0).max() < 1e-11 * scale
    assert np.abs(f[:ns] - f0[:ns]).max() > 1e-3 * scale      # the terms are there
    assert abs(o.En_pot - o0.En_pot) < 1e-11 * abs(o0.En_pot + o0.En_el)
    for k in range(6):
        assert abs(o.Pvir[k] - o0.Pvir[k]) < 1e-10 * abs(o0.Pvir[0])


def test_forces_are_the_energy_gradient():
    # the differences to the kappa = 0 system isolate the shell terms (the
    # erfc splines of energy and force are separate fits, consistent only
    # to about 1e-6)
    r, bs, _ = shellcore.system()
    _, bs0, _ = shellcore.system(kappa=0.0)
    ns = bs.struct.ns
    df = oracle.forces(bs, r)[0] - oracle.forces(bs0, r)[0]
    st = bs.info["site_type"]
    types = ["H", "O", "NA", "CL"]
    h = 1e-5
    checked = 0
    for name in ("CL", "O"):
        for i in np.flatnonzero(st == types.index(name))[:2]:
            for row in (ns + i, i):                    # Drude offset, centre
                ref = df[row] + (df[ns + i] if row == i else 0.0)
                for k in range(3):
                    rp, rm = r.copy(), r.copy()
                    rp[row, k] += h
                    rm[row, k] -= h
                    g = ((energy(bs, rp) - energy(bs0, rp)) - (energy(bs, rm) - energy(bs0, rm))) / (2 * h)
                    print(name, "row", row, "k", k, "-dE/dx %.10e force %.10e" % (-g, ref[k]))
                    assert abs(-g - ref[k]) < 1e-6 * np.abs(ref).max()
                    checked += 1
    assert checked == 24 and np.abs(df).max() > 0


def test_shell_needs_drudes_and_exponential_repulsion():
    _, meta = oracle.load_fixture("brine506")
    ff = shellcore.buck_forcefield(meta["forcefield"])
    with pytest.raises(ValueError):
        build_system(ff, meta["counts"], [meta["L"]] * 3, meta["cutoff"],
                     provider=oracle.provider(), shell={"CL": (-100.0, True)})
    _, pmeta = oracle.load_fixture("polar306")
    with pytest.raises(ValueError):       # LJ form
        build_system(pmeta["forcefield"], pmeta["counts"], [pmeta["L"]] * 3, pmeta["cutoff"],
                     provider=oracle.provider(), shell={"CL": (-100.0, True)})
