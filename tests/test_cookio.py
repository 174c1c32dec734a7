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


def write_deck(td, meta, species, types, polar, nbfix=()):
    ble = os.path.join(td, "sys.ble")
    mg.write_ble(ble, species, types, polar, nbfix)
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
    (("nnbfixes=0", "nnbfixes=1"), "missing data"),        # an nbfix row announced, none given
    (("nparms=0", "nparms=1"), "site-site form"),
    (("nparms=0", "nparms=2"), "Lennard-Jones"),
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


def exp6_ble(tmp_path, a_of, table="exp-6", comb_rule=1):
    """the SPC/E .ble of the fixtures turned into an exp-6 file: nparms=1,
    the exp-6 table, a stiffness column after RvdW"""
    species, types = [mg.spce_species()], ["H", "O"]
    ble = str(tmp_path / "w6.ble")
    mg.write_ble(ble, species, types)
    lines = open(ble).read().replace("nparms=0", "nparms=1").replace(
        "comb_rule=1", "comb_rule=%d" % comb_rule).splitlines()
    k = lines.index("Lennard-Jones")
    lines[k] = table
    for i in range(k + 2, k + 2 + len(types)):
        name = lines[i].split()[1]
        lines[i] += "  %.3f" % a_of[name]
    open(ble, "w").write("\n".join(lines) + "\n")
    return ble, species, types


def test_exp6_force_field_file(tmp_path):
    from macsimus_amd import abi
    from macsimus_amd.system import EUNIT_KCAL
    a_of = {"H": 14.0, "O": 13.5}
    ble, species, types = exp6_ble(tmp_path, a_of)
    ff = cookio.read_ble(ble)
    assert ff["form"] == "exp6" and ff["type_parm"] == [14.0, 13.5]
    assert cookio.read_ble(ble, form="exp6") == ff
    # a deck: cook's double kcal conversion by default, single on request
    open(tmp_path / "w.def", "w").write("N[0]=64 L[0]=12.4 L[1]=12.4 L[2]=12.4 cutoff=6 "
                                        "el.epsr=0.05 el.epsk=0.05;\n")
    prov = oracle.provider()
    bs2, _ = cookio.system_from_cook(ble, str(tmp_path / "w.def"), provider=prov)
    bs1, _ = cookio.system_from_cook(ble, str(tmp_path / "w.def"), provider=prov,
                                     exp6_cook_units=False)
    t2, t1 = bs2.info["sstab"][3], bs1.info["sstab"][3]          # O-O
    assert isinstance(t1, abi.Exp6) and bs1.struct.sstab_exp6
    E_O = -ff["type_EvdW_kcal"][1] * EUNIT_KCAL
    assert t1.eu == pytest.approx(6 * E_O / (13.5 - 6) * np.exp(13.5), rel=1e-12)
    assert t2.eu == pytest.approx(EUNIT_KCAL * t1.eu, rel=1e-12)
    # without a polar key the table is site-types: the form must be named
    ble2, _, _ = exp6_ble(tmp_path, a_of, table="site-types")
    txt = open(ble2).read().replace("polar=0", "")
    open(ble2, "w").write(txt)
    with pytest.raises(cookio.CookInputError, match="site-site form"):
        cookio.read_ble(ble2)
    assert cookio.read_ble(ble2, form="exp6")["type_parm"] == [14.0, 13.5]
    ble3, _, _ = exp6_ble(tmp_path, a_of, comb_rule=0)
    with pytest.raises(cookio.CookInputError, match="Kirkwood"):
        cookio.read_ble(ble3)


def test_eunit_of_the_file(tmp_path):
    """energies given in K with eunit=1 read as the same force field as
    the kcal/mol file (cook/simdef.c:43,531,626)"""
    from macsimus_amd.system import EUNIT_KCAL
    species, types = [mg.spce_species()], ["H", "O"]
    ble = str(tmp_path / "k.ble")
    mg.write_ble(ble, species, types)
    ref = cookio.read_ble(ble)
    lines = open(ble).read().replace("nparms=0;", "nparms=0 eunit=1;").splitlines()
    k = lines.index("Lennard-Jones")
    for i in range(k + 2, k + 2 + len(types)):
        f = lines[i].split()
        f[4] = "%.15g" % (float(f[4]) * EUNIT_KCAL)
        lines[i] = "  " + "  ".join(f)
    open(ble, "w").write("\n".join(lines) + "\n")
    got = cookio.read_ble(ble)
    np.testing.assert_allclose(got["type_EvdW_kcal"], ref["type_EvdW_kcal"], rtol=1e-14)


def test_buckingham_force_field_file(tmp_path):
    """the Buckingham table: A in the EvdW column (positive, not flipped),
    B in RvdW, C in the parameter column; energies scaled by eunit"""
    a_of = {"H": 0.0, "O": 1.2e6}
    ble, species, types = exp6_ble(tmp_path, a_of, table="Buckingham", comb_rule=5)
    txt = open(ble).read()
    lines = txt.splitlines()
    k = lines.index("Buckingham")
    for i in range(k + 2, k + 4):
        f = lines[i].split()
        f[4] = "%.6f" % abs(float(f[4]) * 1e5)           # A_ii > 0
        f[5] = "3.5"                                      # B_ii
        lines[i] = "  " + "  ".join(f)
    open(ble, "w").write("\n".join(lines) + "\n")
    ff = cookio.read_ble(ble)
    assert ff["form"] == "buck" and ff["comb_rule"] == 5
    assert ff["type_EvdW_kcal"][1] > 0 and ff["type_RvdW"][1] == 3.5
    assert ff["type_parm"] == [0.0, 1.2e6]


def test_busing_force_field_file(tmp_path):
    """the Busing table: f rho_i in the EvdW column (kept as written, a
    negative one switches the site off), R_i in RvdW, C_i in the parameter
    column; EvdW scaled by eunit, C_i by its square root (Epow 1 and 1/2,
    sim/busing/sitesite.c:5-16); the deck builds the Busing table"""
    from macsimus_amd.system import EUNIT_KCAL
    a_of = {"H": 5.0, "O": 40.0}
    ble, species, types = exp6_ble(tmp_path, a_of, table="Busing", comb_rule=0)
    lines = open(ble).read().splitlines()
    k = lines.index("Busing")
    for i, (E, R) in zip(range(k + 2, k + 4), ((-1.0, 0.0), (6.2, 1.8))):
        f = lines[i].split()
        f[4], f[5] = "%.3f" % E, "%.3f" % R
        lines[i] = "  " + "  ".join(f)
    open(ble, "w").write("\n".join(lines) + "\n")
    ff = cookio.read_ble(ble)
    assert ff["form"] == "busing" and ff["comb_rule"] == 0
    assert ff["type_EvdW_kcal"] == [-1.0, 6.2] and ff["type_RvdW"] == [0.0, 1.8]
    assert ff["type_parm"] == [5.0, 40.0]
    txt = open(ble).read().replace("comb_rule=0", "comb_rule=0 eunit=%.15g" % (2 * EUNIT_KCAL))
    open(ble, "w").write(txt)
    ff2 = cookio.read_ble(ble)
    assert ff2["type_EvdW_kcal"] == pytest.approx([-2.0, 12.4], rel=1e-14)
    assert ff2["type_parm"] == pytest.approx([5.0 * 2 ** 0.5, 40.0 * 2 ** 0.5], rel=1e-14)
    open(tmp_path / "w.def", "w").write("N[0]=64 L[0]=12.4 L[1]=12.4 L[2]=12.4 cutoff=6 "
                                        "el.epsr=0.05 el.epsk=0.05;\n")
    bs, _ = cookio.system_from_cook(ble, str(tmp_path / "w.def"), provider=oracle.provider())
    t = bs.info["sstab"]
    assert bs.struct.sstab_exp6 and t[0].eu == 0 and t[1].eu == 0
    assert t[3].eu * np.exp(t[3].tau * 3.6) == pytest.approx(2 * 12.4 * EUNIT_KCAL, rel=1e-13)
    assert t[3].ru == pytest.approx((40.0 * 2 ** 0.5) ** 2 * EUNIT_KCAL, rel=1e-13)


def test_nbfix_table(tmp_path):
    """nbfixes (cook/simdef.c:772-798, sim/setss.c:190-199): the pair's own
    Emin and RvdW replace the combining rule for that type pair.  An O-O
    nbfix of the values the rule gives (Emin_O, 2 RvdW_O) leaves the table
    and the oracle's forces as they were; other values give cook's LJ
    constants of initssaux() (sim/lj/sitesite.c:96-104): E4 = -4 Emin,
    Sq = RvdW^2 2^(-1/3); other pairs keep the rule."""
    from macsimus_amd.system import EUNIT_KCAL
    species, types = [mg.spce_species()], ["H", "O"]
    ble = str(tmp_path / "fix.ble")
    mg.write_ble(ble, species, types)
    base = cookio.read_ble(ble)
    E_O, R_O = base["type_EvdW_kcal"][1], base["type_RvdW"][1]
    open(tmp_path / "w.def", "w").write("N[0]=64 L[0]=12.4 L[1]=12.4 L[2]=12.4 cutoff=6 "
                                        "el.epsr=0.05 el.epsk=0.05;\n")
    prov = oracle.provider()
    bs0, _ = cookio.system_from_cook(ble, str(tmp_path / "w.def"), provider=prov)
    txt = open(ble).read().replace("nnbfixes=0", "nnbfixes=1")

    def with_row(row):
        open(ble, "w").write(txt.replace("! atom atom Eref R\n", "! atom atom Eref R\n%s\n" % row))
        return cookio.system_from_cook(ble, str(tmp_path / "w.def"), provider=prov)[0]

    bs1 = with_row("O O %.15g %.15g" % (E_O, 2 * R_O))
    assert cookio.read_ble(ble)["nbfix"] == [["O", "O", E_O, 2 * R_O]]
    t0, t1 = bs0.info["sstab"][3], bs1.info["sstab"][3]
    for k in ("Sq", "E4", "E24", "E48", "C1q", "C2q", "A", "A4"):
        assert getattr(t1, k) == pytest.approx(getattr(t0, k), rel=1e-13), k
    rng = np.random.default_rng(3)
    r = rng.uniform(0, 12.4, (192, 3))
    f0, o0 = oracle.forces(bs0, r, 1)
    f1, o1 = oracle.forces(bs1, r, 1)
    assert np.abs(f1 - f0).max() < 1e-9 * np.abs(f0).max()
    bs2 = with_row("O O -0.2 3.4")                  # other values
    t2 = bs2.info["sstab"][3]
    assert t2.E4 == pytest.approx(0.8 * EUNIT_KCAL, rel=1e-13)
    assert t2.Sq == pytest.approx(3.4 ** 2 * 2 ** (-1 / 3), rel=1e-13)
    for k in (0, 1, 2):
        assert bs2.info["sstab"][k].E4 == bs0.info["sstab"][k].E4
    # types by name only (atomn(), cook/simdef.c:77-120): a type number is
    # not a type
    with pytest.raises(cookio.CookInputError):
        with_row("1 1 -0.2 3.4")
    # the exp-6 form refuses nbfixes
    ble6, _, _ = exp6_ble(tmp_path, {"H": 14.0, "O": 13.5})
    open(ble6, "w").write(open(ble6).read().replace("nnbfixes=0", "nnbfixes=1"))
    with pytest.raises(cookio.CookInputError):
        cookio.read_ble(ble6)


def test_shell_core_columns(tmp_path):
    """the kappa and rep columns of a polar file (cook/simdef.c:646-694):
    read into ff["shell"]; applied only with shell_core=True (cook's POLAR&1
    build), which needs the Buckingham or Busing form"""
    d, meta = oracle.load_fixture("polar306")
    species, types, _ = CASES["polar306"]()
    polar = {k: tuple(v) for k, v in meta["forcefield"]["polar"].items()}
    ble, dfn = write_deck(str(tmp_path), meta, species, types, polar)
    lines = open(ble).read().splitlines()
    for i, ln in enumerate(lines):
        f = ln.split()
        if len(f) == 10 and f[1] in ("CL", "O", "NA"):
            f[9] = {"CL": "-245.27", "O": "-100", "NA": "0"}[f[1]]   # kappa
            f.append("1")                         # rep
            lines[i] = "  " + "  ".join(f)
    open(ble, "w").write("\n".join(lines) + "\n")
    ff = cookio.read_ble(ble)
    assert ff["shell"] == {"O": (-100.0, 1), "NA": (0.0, 1), "CL": (-245.27, 1)}
    bs, _ = cookio.system_from_cook(ble, dfn, provider=oracle.provider())
    assert not bs.struct.shell_qkappa
    with pytest.raises(ValueError, match="Buckingham or Busing"):
        cookio.system_from_cook(ble, dfn, provider=oracle.provider(), shell_core=True)


def test_nbfix_deck_gives_cooks_forces(tmp_path):
    """cook's own run of the brine deck with an Na-O nbfix (tests/golden/
    brine506_nbfix: Emin 0.2 kcal/mol at 3.0 A): the deck read by cookio
    gives, through the oracle, cook's forces, Epot and tensor; without the
    nbfix row the forces differ"""
    d, meta = oracle.load_fixture("brine506_nbfix")
    assert meta["nbfix"] == [["NA", "O", -0.2, 3.0]]
    species, types, _ = CASES["brine506"]()
    ble, dfn = write_deck(str(tmp_path), meta, species, types, None,
                          [tuple(x) for x in meta["nbfix"]])
    assert cookio.read_ble(ble)["nbfix"][0][:2] == ["NA", "O"]
    bs, _ = cookio.system_from_cook(ble, dfn, provider=oracle.provider())
    f, out = oracle.forces(bs, d["r"], 1)
    err = float(np.sqrt(((f - d["f"]) ** 2).sum() / (d["f"] ** 2).sum()))
    assert err < 1e-10, err
    assert abs(out.En_pot + out.En_el - meta["Epot"]) < 1e-9 * abs(meta["Epot"])
    ble0, dfn0 = write_deck(str(tmp_path / "plain"), meta, species, types, None) \
        if (tmp_path / "plain").mkdir() is None else (None, None)
    bs0, _ = cookio.system_from_cook(ble0, dfn0, provider=oracle.provider())
    f0, _ = oracle.forces(bs0, d["r"], 1)
    assert float(np.sqrt(((f0 - d["f"]) ** 2).sum() / (d["f"] ** 2).sum())) > 1e-5
