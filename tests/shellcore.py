"""The shell-core (POLAR&1) test system: cook's polar306 configuration
(300 Drude water + 3 Drude NaCl, Drudes on O and Cl) with the Buckingham
form of each site's exp-6 parameters (A, B, C of the identical-site
mapping, geometric rules), shells (qkappa) on the Drude types O and Cl and
the rep flag on O, Na and Cl."""
import math

import numpy as np

import oracle
from macsimus_amd import workloads
from macsimus_amd.system import ELECTRON


def buck_forcefield(ff):
    ff = dict(ff, form="buck", comb_rule=5)
    A, B, Cc = [], [], []
    for t, E, R in zip(ff["types"], ff["type_EvdW_kcal"], ff["type_RvdW"]):
        a, sig, eps = workloads.EXP6_A[t], 2 * R, -E
        A.append(6 * eps * math.exp(a) / (a - 6))
        B.append(a / sig)
        Cc.append(eps * a * sig ** 6 / (a - 6))
    ff.update(type_EvdW_kcal=A, type_RvdW=B, type_parm=Cc)
    return ff


def system(kappa=-0.6, provider_=None, shells=("O", "CL"), rep=("O", "NA", "CL")):
    """(positions [2 ns][3] with the Drudes of the fixture's first SCF
    iteration, built system, fixture meta); kappa is si->qkappa, the
    factor of the repulsion (cook's ATOM kappa = kappa * electron)"""
    d, meta = oracle.load_fixture("polar306")
    ff = buck_forcefield(meta["forcefield"])
    shell = {t: ((kappa if t in shells else 0.0) * ELECTRON, t in rep) for t in ff["types"]}
    bs = oracle.fixture_system(dict(meta, forcefield=ff), provider_=provider_, shell=shell)
    r = np.concatenate([d["r"], d["rpol"][1]])
    return r, bs, meta
