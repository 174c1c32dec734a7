"""Ewald corrections of cook's Ewald(2) beyond the self and dipole terms
(sim/ewald.c:477-492, 865-997; sim/ewaldprt.c:13): the background energy
of a charged box, the Yeh-Berkowitz slab correction with and without the
charged-background force, and the diagonal term.

The oracle matches cook's fixtures for these decks (test_oracle_golden.py).
Here each term is shown to matter at the fixtures' precision: switching
it off in the oracle moves forces or energy away from cook's by far more
than the agreement, so a dropped or mis-signed term fails.  The engine's
set-up rejects what cook's set-up rejects (host checks, no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from macsimus_amd import abi
from macsimus_amd.system import engine_provider


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def run(name, **kw):
    d, meta = oracle.load_fixture(name)
    args = dict(meta.get("ewald", {}))
    args.update(kw)
    bs = oracle.fixture_system(meta, **args)
    f, out = oracle.forces(bs, d["r"])
    return d, meta, bs, f, out


def test_background_energy_of_a_charged_box():
    d, meta, bs, f, out = run("brine505_charged")
    E = out.En_pot + out.En_el
    assert abs(E - meta["Epot"]) < 1e-11 * abs(meta["Epot"])
    # -pi/2 (Q/alpha)^2 / V with Q = +1 e
    Q = float(bs.info["charge"].sum())
    assert abs(Q - 408.779826922048) < 1e-6
    bg = -np.pi / 2 * (Q / bs.info["alpha"]) ** 2 / meta["L"] ** 3
    assert abs(bg) > 1e5 * 1e-11 * abs(meta["Epot"])
    # without the term the energy misses cook's by exactly bg
    assert abs((E - bg) - meta["Epot"]) > 0.99 * abs(bg)


def test_slab_correction_changes_forces_and_tensor():
    d, meta, bs, f, out = run("brine506_slab")
    assert rel_rms(f, d["f"]) < 1e-10
    _, _, _, f0, out0 = run("brine506_slab", el_corr=0)
    assert rel_rms(f0, d["f"]) > 1e3 * 1e-10
    # the slab term is the z dipole only: x and y forces are the tin-foil ones
    assert rel_rms(f0[:, :2], f[:, :2]) < 1e-12
    # energy only (el.corr + 8): the energy is the slab one, the forces the
    # tin-foil ones
    _, _, _, f8, out8 = run("brine506_slab", el_corr=12)
    assert abs((out8.En_el + out8.En_pot) - meta["Epot"]) < 1e-11 * abs(meta["Epot"])
    assert rel_rms(f8, f0) < 1e-12


def test_slab_background_force_and_diagonal_term():
    d, meta, bs, f, out = run("brine505_slabbg")
    assert rel_rms(f, d["f"]) < 1e-10
    E = out.En_pot + out.En_el
    assert abs(E - meta["Epot"]) < 1e-11 * abs(meta["Epot"])
    # Ediag: cook adds it for el.diag = 1
    _, _, _, fd, outd = run("brine505_slabbg", el_diag=0)
    dE = (out.En_el - outd.En_el)
    assert abs(dE) > 1e3 * 1e-11 * abs(meta["Epot"])
    assert rel_rms(fd, f) < 1e-14
    # Ediag from its definition: the k lattice between kappa and
    # sqrt(kappa^2 + alpha^2), summed here with numpy
    L, a, k = meta["L"], bs.info["alpha"], bs.info["kappa"]
    K = int(np.sqrt(k * k + a * a) * L)
    g = np.arange(-K, K + 1) / L
    kk = (g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2).ravel()
    sel = (kk > k * k) & (kk < k * k + a * a)
    Ed = np.exp(-(np.pi / a) ** 2 * kk[sel]).dot(1 / kk[sel]) * bs.info["sumq"] / (2 * np.pi * L ** 3)
    assert abs(dE - Ed) < 1e-9 * abs(Ed)


def test_background_force_restores_translation_invariance():
    """For a charged box the Yeh-Berkowitz field -c M_k moves by -c Q t
    when every site moves by t along k; el.bg's field q 4 pi Q r_k / V
    moves by exactly the opposite, so the forces with el.bg do not change
    (Ballenegger et al., the reason for the term) and those without it
    shift by -q c Q t (the oracle takes el_bg = 0 for a charged box; the
    engine, like cook, does not)"""
    d, meta = oracle.load_fixture("brine505_slabbg")
    t = 0.37
    r2 = d["r"].copy()
    r2[:, 0] += t
    bs = oracle.fixture_system(meta)
    f1, _ = oracle.forces(bs, d["r"])
    f2, _ = oracle.forces(bs, r2)
    assert rel_rms(f2, f1) < 1e-10
    bs0 = oracle.fixture_system(meta, el_bg=0)
    g1, _ = oracle.forces(bs0, d["r"])
    g2, _ = oracle.forces(bs0, r2)
    q = bs.info["charge"]
    Q = float(q.sum())
    exp = np.zeros_like(g1)
    exp[:, 0] = -q * (4 * np.pi / meta["L"] ** 3) * Q * t
    assert rel_rms(g2 - g1, exp) < 1e-6
    assert rel_rms(g1, f1) > 1e-6


def _system_with(**kw):
    d, meta = oracle.load_fixture("brine506")
    return oracle.fixture_system(meta, provider_=oracle.provider(), **kw), meta


def _create(bs):
    lib = abi.load_engine()
    ctx = C.c_void_p()
    rc = lib.mdx_create(C.byref(ctx), C.byref(bs.struct), abi.MDX_MIXED)
    msg = (lib.mdx_last_error() or b"").decode()
    if rc == 0:
        lib.mdx_destroy(ctx)
    return rc, msg


def test_create_rejects_what_cook_rejects():
    """host checks of mdx_create, made before any device call"""
    d, meta = oracle.load_fixture("brine505_charged")
    bs = oracle.fixture_system(meta, provider_=oracle.provider(), el_bg=0)
    rc, msg = _create(bs)
    assert rc == -1 and "el_bg" in msg
    bs = oracle.fixture_system(meta, provider_=oracle.provider(), el_bg=1, epsinf=80.0)
    rc, msg = _create(bs)
    assert rc == -1 and "tin-foil" in msg
    bs, _ = _system_with(el_corr=4, epsinf=1e10)
    rc, msg = _create(bs)
    assert rc == -1 and "Yeh-Berkowitz" in msg
    bs, _ = _system_with(el_corr=16)
    rc, msg = _create(bs)
    assert rc == -1 and "el_corr" in msg
    # a site type outside the table
    bs, _ = _system_with()
    st = np.ctypeslib.as_array(bs.struct.site_type, shape=(bs.struct.ns,))
    st[7] = bs.struct.ntypes
    rc, msg = _create(bs)
    assert rc == -1 and "site type" in msg
    st[7] = -1
    rc, msg = _create(bs)
    assert rc == -1 and "site type" in msg
