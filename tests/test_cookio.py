"""cook's input files drive the engine: the .ble force field and .def deck
that the golden fixtures were made from (written by the fixture script's
own writers) are read back by macsimus_amd.cookio, give the same system
arrays as the fixtures' stored force field, and the oracle's forces on
them are cook's; malformed or unsupported input is rejected."""
import os
import sys

import numpy as np
import pytest

import oracle
from macsimus_amd import cookio

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as mg  # noqa: E402  (the writers only; nothing is run)

CASES = {
    # fixture: (species, types, polar) as make_golden builds them
    "spce1000_lattice": lambda: ([mg.spce_species()], ["H", "O"], None),
    "brine506": lambda: ([mg.ion_species("NA", 1.0), mg.ion_species("CL", -1.0), mg.spce_species()],
                         ["H", "O", "NA", "CL"], None),
    "polar306": lambda: ([mg.ion_species("NA", 1.0), mg.ion_species("CL", -1.0), mg.spce_species()],
                         ["H", "O", "NA", "CL"], None),
}


def write_deck(td, meta, species, types, polar):
    ble = os.path.join(td, "sys.ble")
    mg.write_ble(ble, species, types, polar)
    L = meta["L"]
    lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(meta["counts"])]
    lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L), "cutoff=%g" % meta["cutoff"],
              "LJcutoff=cutoff", "el.epsr=%g" % meta["epsr"], "el.epsk=%g" % meta["epsk"],
              "h=0.002 noint=1 T=298 epsc=1e-10", ";"]
    dfn = os.path.join(td, "sys.def")
    open(dfn, "w").write("\n".join(lines) + "\n")
    return ble, dfn


def arrays(bs):
    S = bs.struct
    i = bs.info
    return dict(site_type=i["site_type"], charge=i["charge"], chargepol=i["chargepol"],
                excl=i["excl"], mass=i["mass"], cons_ij=i["cons_ij"], cons_bondq=i["cons_bondq"],
                sstab=np.array([[getattr(S.sstab[k], f) for f in ("Sq", "E4", "E24", "E48", "C1q",
                                                                  "C2q", "A", "A4")]
                                for k in range(S.ntypes ** 2)]),
                alpha=S.alpha, kappa=S.kappa, cutoff=S.cutoff, L=list(S.L))


@pytest.mark.parametrize("name", ["spce1000_lattice", "brine506", "polar306"])
def test_cook_inputs_give_the_fixture_system(name, tmp_path):
    d, meta = oracle.load_fixture(name)
    species, types, _ = CASES[name]()
    polar = meta["forcefield"].get("polar")
    polar = {k: tuple(v) for k, v in polar.items()} if polar else None
    ble, dfn = write_deck(str(tmp_path), meta, species, types, polar)
    ff = cookio.read_ble(ble)
    assert ff["types"] == types
    assert [sp["name"] for sp in ff["species"]] == [sp["name"] for sp in species]
    ref = meta["forcefield"]
    for k in ("type_mass", "type_EvdW_kcal", "type_RvdW"):
        assert np.allclose(ff[k], ref[k], rtol=1e-11, atol=1e-12), k
    for sp, rsp in zip(ff["species"], ref["species"]):
        assert sp["site_type"] == rsp["site_type"] and sp["site_excl"] == rsp["site_excl"]
        assert np.allclose(sp["site_charge_e"], rsp["site_charge_e"], atol=1e-12)
        assert np.allclose(np.array(sp["bonds"]), np.array(rsp["bonds"]), atol=1e-11)
    bs, params = cookio.system_from_cook(ble, dfn, provider=oracle.provider())
    assert params["LJcutoff"] == meta["cutoff"]
    ref_bs = oracle.fixture_system(meta)
    a, b = arrays(bs), arrays(ref_bs)
    for k in a:
        assert np.allclose(np.asarray(a[k], float), np.asarray(b[k], float), rtol=1e-10, atol=1e-12), k
    # the oracle on the system read from cook's files gives cook's forces
    # (Drude fixtures: centres then the Drude offsets of the first dump)
    if "rpol" in d:
        r = np.concatenate([d["r"], d["rpol"][0]])
        ref_f = np.concatenate([d["f"][0], d["fpol"][0]])
    else:
        r, ref_f = d["r"], d["f"]
    f, out = oracle.forces(bs, r, 1)
    e = np.sqrt(((f - ref_f) ** 2).sum() / (ref_f ** 2).sum())
    assert e < 1e-9, e


def test_getdata_expressions_and_strings():
    d, rest = cookio.parse_getdata('! comment = 1\nN[0]=10 L[0]=2.5 cutoff=1 LJcutoff=cutoff\n'
                                   'a=(2*cutoff+L[0])/2 el.epsr=1e-5 init="crystal" ;tail')
    assert d == {"N[0]": 10, "L[0]": 2.5, "cutoff": 1, "LJcutoff": 1, "a": 2.25, "el.epsr": 1e-5,
                 "init": "crystal"}
    assert rest == "tail"
    with pytest.raises(cookio.CookInputError):
        cookio.parse_getdata("x=undefined_name*2;")
    with pytest.raises(cookio.CookInputError):
        cookio.parse_getdata("x=__import__;")


@pytest.mark.parametrize("change,msg", [
    (("nnbfixes=0", "nnbfixes=1"), "nbfixes"),
    (("nparms=0", "nparms=1"), "Lennard-Jones form"),
])
def test_unsupported_input_is_rejected(tmp_path, change, msg):
    species, types = [mg.spce_species()], ["H", "O"]
    ble = str(tmp_path / "w.ble")
    mg.write_ble(ble, species, types)
    text = open(ble).read()
    assert change[0] in text
    open(ble, "w").write(text.replace(change[0], change[1]))
    with pytest.raises(cookio.CookInputError, match=msg):
        cookio.read_ble(ble)


def test_one_four_and_vibrating_bonds_are_rejected(tmp_path):
    species, types = [mg.spce_species()], ["H", "O"]
    ble = str(tmp_path / "w.ble")
    mg.write_ble(ble, species, types)
    text = open(ble).read()
    # a 1-4 exception on the second H, then a bond force constant
    lines = text.splitlines()
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("1 H") and "0.4238" in l)
    bad = lines[:]
    bad[k] = bad[k].rsplit(" ", 1)[0] + " *0"
    open(ble, "w").write("\n".join(bad) + "\n")
    with pytest.raises(cookio.CookInputError, match="1-4"):
        cookio.read_ble(ble)
    lines2 = text.splitlines()
    j = next(i for i, l in enumerate(lines2) if l.strip().startswith("0 H    2 O"))
    parts = lines2[j].split()
    parts[4] = "450"
    lines2[j] = " " + " ".join(parts)
    open(ble, "w").write("\n".join(lines2) + "\n")
    with pytest.raises(cookio.CookInputError, match="vibrating"):
        cookio.read_ble(ble)
