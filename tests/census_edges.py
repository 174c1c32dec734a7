"""Configurations at the edges of cook's linked-cell census (test helper).

cook decides a pair by rr = |(r_c + shift) - r_n|^2 < cutoff^2 in its own
arithmetic (cook/lc.c:217-420, cook/ljqqnp.c:74-85): the central site's
local copy is shifted by +-L where the neighbour cell wraps, coordinates
are wrapped by subtraction and negative ones stay (cook/linklist.c:
155-176), cells use a shaved rL (cook/linklist.c:81-86).  A pair whose
rr is within a few ulps of cutoff^2 can be decided differently by any
other arithmetic of the same distance.  The search below finds two-ion
configurations across a periodic boundary where the oracle (the
restatement of cook's traversal, pinned on cook's fixtures) and a plain
minimum-image evaluation r_i - (r_j + shift) disagree, so that a census
not built on cook's arithmetic fails on them.
"""
import json
import os

import numpy as np

import oracle

L = 30.0
CUTOFF = 9.0
NCELL = (6, 6, 6)


def ion_ff():
    """Na+ and Cl- of the brine fixtures (LJ + point charge)"""
    ff = json.load(open(os.path.join(oracle.GOLDEN, "brine506.json")))["forcefield"]
    return ff


def ion_system(n_na, n_cl, provider_=None):
    return oracle.build_system(ion_ff(), counts=[n_na, n_cl, 0], L=[L] * 3, cutoff=CUTOFF,
                               ncell=NCELL, provider=provider_ or oracle.provider())


def naive_npair(r):
    """census by the plain minimum image r_i - (r_j + shift) of coordinates
    wrapped into [0, L) by floor (both orders, as a full-shell kernel)"""
    w = r - L * np.floor(r / L)
    n = 0.0
    for i in range(len(w)):
        for j in range(len(w)):
            if i == j:
                continue
            d = w[j] - w[i]
            s = L * np.rint(d / L)
            dd = w[i] - (w[j] - s)
            n += 0.5 * float((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2] < CUTOFF * CUTOFF)
    return n


def sharp_pairs(n, seed=3, axis=0):
    """n two-ion configurations (Na at row 0, Cl at row 1) whose census by
    the oracle differs from naive_npair: the pair lies across the periodic
    boundary along `axis` at a distance within a few ulps of the cutoff"""
    bs = ion_system(1, 1)
    rng = np.random.default_rng(seed)
    cq = CUTOFF * CUTOFF
    out = []
    for _ in range(4000):
        a = rng.uniform(0.5, L - 0.5, 3)
        a[axis] = rng.uniform(0.0, 1.5)
        dv = rng.normal(size=3)
        dv[axis] = 0
        dv *= rng.uniform(0, 5.0) / np.linalg.norm(dv)
        dx = np.sqrt(cq - (dv ** 2).sum())
        b = a + dv
        b[axis] = a[axis] - dx + L
        b = np.where(b >= L, b - L, b)
        b = np.where(b < 0, b + L, b)
        x0 = b[axis]
        for k in range(-12, 13):
            b[axis] = x0 + k * np.spacing(x0)
            r = np.array([a, b])
            _, o = oracle.forces(bs, r, measure=0, nthreads=1)
            if o.npair != naive_npair(r):
                out.append(r.copy())
                break
        if len(out) == n:
            break
    return out


def edge_sites():
    """12 ions at the edges of cook's cell rule along x: 0, -1e-12, L - ulp,
    L, exactly on k L/nc, just below and above them, and beyond L (wrapped
    by subtraction); y and z keep them apart"""
    nc = NCELL[0]
    xs = [0.0, -1e-12, L - np.spacing(L), L, 1 * L / nc, 2 * L / nc, np.nextafter(3 * L / nc, 0),
          np.nextafter(4 * L / nc, L), 5 * L / nc, L + 0.25, 2 * L - np.spacing(2 * L), 1e-300]
    r = np.zeros((len(xs), 3))
    for i, x in enumerate(xs):
        r[i] = (x, 2.5 * i, (7.3 * i) % L)
    return r
