"""ctypes access to the CPU FP64 restatement in oracle/ (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module; the engine never does.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

from macsimus_amd import abi
from macsimus_amd.system import Provider, build_system

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "libmdx_oracle.so")
GOLDEN = os.path.join(REPO, "tests", "golden")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
        lib = C.CDLL(ORACLE_SO)
        abi.declare_setup(lib, "oracle_")
        dp = C.POINTER(C.c_double)
        lib.oracle_forces.argtypes = [C.POINTER(abi.System), dp, dp,
                                      C.POINTER(abi.Out), C.c_int, C.c_int,
                                      C.POINTER(C.c_int)]
        lib.oracle_forces.restype = C.c_int
        lib.oracle_md_step.argtypes = [C.POINTER(abi.System), dp, dp, dp,
                                       C.POINTER(abi.Step), C.c_int,
                                       C.POINTER(abi.StepOut), C.c_int]
        lib.oracle_md_step.restype = C.c_int
        _lib = lib
    return _lib


def provider():
    return Provider(load(), "oracle_")


def load_fixture(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    return d, meta


def fixture_system(meta, provider_=None, **kw):
    L = [meta["L"]] * 3
    args = dict(counts=meta["counts"], L=L, cutoff=meta["cutoff"],
                epsr=meta["epsr"], epsk=meta["epsk"],
                ncell=tuple(meta.get("ncell", (0, 0, 0))))
    args.update(meta.get("ewald", {}))
    args.update(kw)
    return build_system(meta["forcefield"], provider=provider_ or provider(), **args)


def forces(bs, r, measure=1, nthreads=None, cellid=False):
    lib = load()
    r = np.ascontiguousarray(r, dtype=np.float64)
    f = np.zeros_like(r)
    out = abi.Out()
    nt = nthreads or os.cpu_count() or 1
    cid = np.zeros(bs.struct.ns, np.int32) if cellid else None
    rc = lib.oracle_forces(C.byref(bs.struct),
                           r.ctypes.data_as(C.POINTER(C.c_double)),
                           f.ctypes.data_as(C.POINTER(C.c_double)),
                           C.byref(out), measure, nt,
                           cid.ctypes.data_as(C.POINTER(C.c_int)) if cellid else None)
    if rc:
        raise RuntimeError("oracle_forces failed (%d)" % rc)
    return (f, out, cid) if cellid else (f, out)


def rdf_shape(ntypes, grid, cutoff):
    """(records, bins) of cook's RDF histograms (sim/simmeas.c:655-767)"""
    g = int(grid)
    nrec = 1 if g < 0 else ntypes * (ntypes + 1) // 2
    return nrec, int(cutoff * abs(g) + 1.0000001)


def forces_rdf(bs, r, grid, cutoff, hist=None, nthreads=None):
    """a measured force call adding its RDF histograms to hist (records of
    type pairs (i, j<=i), uint64); returns (f, out, hist)"""
    lib = load()
    dp = C.POINTER(C.c_double)
    lib.oracle_forces_rdf.argtypes = [C.POINTER(abi.System), dp, dp, C.POINTER(abi.Out),
                                      C.c_int, C.POINTER(C.c_int), C.c_int, C.c_double,
                                      C.POINTER(C.c_ulonglong)]
    r = np.ascontiguousarray(r, dtype=np.float64)
    f = np.zeros_like(r)
    out = abi.Out()
    shape = rdf_shape(bs.struct.ntypes, grid, cutoff)
    if hist is None:
        hist = np.zeros(shape, np.uint64)
    if hist.shape != shape or hist.dtype != np.uint64 or not hist.flags.c_contiguous:
        raise ValueError("hist must be a C-contiguous uint64 array of shape %s" % (shape,))
    rc = lib.oracle_forces_rdf(C.byref(bs.struct), r.ctypes.data_as(dp), f.ctypes.data_as(dp),
                               C.byref(out), nthreads or os.cpu_count() or 1, None, int(grid),
                               cutoff, hist.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    if rc:
        raise RuntimeError("oracle_forces_rdf failed (%d)" % rc)
    return f, out, hist


def selffield(bs, r, f, out, omega):
    """one selffield() iteration on r (2ns rows, Drude offsets updated) and
    f (2ns rows, Drude forces passed to the centres); adds to out."""
    lib = load()
    lib.oracle_selffield.argtypes = [C.POINTER(abi.System), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(abi.Out), C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
    es, me = C.c_double(0), C.c_double(0)
    rc = lib.oracle_selffield(C.byref(bs.struct), r.ctypes.data_as(C.POINTER(C.c_double)),
                              f.ctypes.data_as(C.POINTER(C.c_double)), C.byref(out), omega,
                              C.byref(es), C.byref(me))
    if rc:
        raise RuntimeError("oracle_selffield failed (%d)" % rc)
    return es.value, me.value


def md_step_polar(bs, r, r1, hist, kk, p, measure=1, nthreads=None):
    """one ASPC + SCF + SHAKE step; r, r1 with 2ns rows (Drude rows of r1
    hold the history aa[1]); hist (k, ns, 3) = aa[2..k+1], updated in
    place; returns (StepOut, new kk)"""
    lib = load()
    dp = C.POINTER(C.c_double)
    lib.oracle_md_step_polar.argtypes = [C.POINTER(abi.System), dp, dp, dp, dp,
                                         C.POINTER(C.c_int), C.POINTER(abi.Step), C.c_int,
                                         C.POINTER(abi.StepOut), C.c_int]
    f = np.zeros_like(r)
    out = abi.StepOut()
    k = C.c_int(kk)
    rc = lib.oracle_md_step_polar(C.byref(bs.struct), r.ctypes.data_as(dp), r1.ctypes.data_as(dp),
                                  f.ctypes.data_as(dp), hist.ctypes.data_as(dp), C.byref(k),
                                  C.byref(p), measure, C.byref(out), nthreads or os.cpu_count() or 1)
    if rc:
        raise RuntimeError("oracle_md_step_polar failed (%d)" % rc)
    return out, k.value


def md_step(bs, r, r1, p, measure=1, nthreads=None):
    lib = load()
    f = np.zeros_like(r)
    out = abi.StepOut()
    nt = nthreads or os.cpu_count() or 1
    rc = lib.oracle_md_step(C.byref(bs.struct),
                            r.ctypes.data_as(C.POINTER(C.c_double)),
                            r1.ctypes.data_as(C.POINTER(C.c_double)),
                            f.ctypes.data_as(C.POINTER(C.c_double)),
                            C.byref(p), measure, C.byref(out), nt)
    if rc:
        raise RuntimeError("oracle_md_step failed (%d)" % rc)
    return out


# ---- the cook run with ASPC (tests/golden/polar306_aspc): -p420, k = 4,
# one SCF iteration per step; dump k holds r(t_k) and the predicted Drude
# offsets of step k.
ASPC_K = 4


def aspc_omega(kk):
    """stability value of scf.omega once the predictor reached length kk
    (sim/pred2.c:75)"""
    return (kk + 2) / (2 * kk + 3)


def aspc_post_scf(bs, d):
    """post-SCF Drude offsets of each dumped step: one selffield() from the
    dump with the omega of that step"""
    ns = bs.struct.ns
    post = {}
    for k in range(1, d["r"].shape[0]):
        r = np.concatenate([d["r"][k], d["rpol"][k]])
        f, out = forces(bs, r)
        selffield(bs, r, f, out, aspc_omega(min(k, ASPC_K)))
        post[k] = r[ns:].copy()
    return post


def aspc_state(d, post, k):
    """(r, r1, hist) before step k in cook's layout: Drude rows of r =
    aa[0], of r1 = aa[1], hist = aa[2..k+1]; needs k >= ASPC_K + 3"""
    r = np.concatenate([d["r"][k], post[k - 1]])
    r1 = np.concatenate([d["r"][k] - d["r"][k - 1], post[k - 2]])
    hist = np.stack([post[k - 3 - m] for m in range(ASPC_K)])
    return r, r1, hist


def md_step_npt(bs, r, r1, L, L0, npt, kk, hist, p, nthreads=None):
    """one Shake() step with the isotropic MTK/Hoover barostat; r, r1, the
    box L [3], npt [6] (logs, logs1, lambda1, En.ext, En.Pref, dlogs) and
    hist [k][2+3ns] updated in place; L0 is the box of the k-vector set;
    returns (StepOut, new kk)"""
    lib = load()
    dp = C.POINTER(C.c_double)
    lib.oracle_md_step_npt.argtypes = [C.POINTER(abi.System), dp, dp, dp, dp, dp,
                                       C.POINTER(abi.Step), dp, C.POINTER(C.c_int), dp,
                                       C.POINTER(abi.StepOut), C.c_int]
    for a in (r, r1, L, L0, npt, hist):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    assert hist.shape[0] >= max(p.trvp_k, 1) and hist.shape[1] == 2 + r.size
    f = np.zeros_like(r)
    out = abi.StepOut()
    k = C.c_int(kk)
    rc = lib.oracle_md_step_npt(C.byref(bs.struct), r.ctypes.data_as(dp), r1.ctypes.data_as(dp),
                                f.ctypes.data_as(dp), L.ctypes.data_as(dp), L0.ctypes.data_as(dp),
                                C.byref(p), npt.ctypes.data_as(dp), C.byref(k),
                                hist.ctypes.data_as(dp), C.byref(out),
                                nthreads or os.cpu_count() or 1)
    if rc:
        raise RuntimeError("oracle_md_step_npt failed (%d)" % rc)
    return out, k.value


def md_step_nose(bs, r, r1, nose, kk, hist, p, measure=1, nthreads=None):
    """one Shake() step with Nose-Hoover + TRVP; r, r1, nose [4] (logs,
    logs1, En.ext, dlogs) and hist [k][1+3ns] updated in place; returns
    (StepOut, new kk)"""
    lib = load()
    dp = C.POINTER(C.c_double)
    lib.oracle_md_step_nose.argtypes = [C.POINTER(abi.System), dp, dp, dp, C.POINTER(abi.Step),
                                        dp, C.POINTER(C.c_int), dp, C.c_int,
                                        C.POINTER(abi.StepOut), C.c_int]
    for a in (r, r1, nose, hist):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    assert hist.shape[0] >= max(p.trvp_k, 1) and hist.shape[1] == 1 + r.size
    f = np.zeros_like(r)
    out = abi.StepOut()
    k = C.c_int(kk)
    rc = lib.oracle_md_step_nose(C.byref(bs.struct), r.ctypes.data_as(dp), r1.ctypes.data_as(dp),
                                 f.ctypes.data_as(dp), C.byref(p), nose.ctypes.data_as(dp),
                                 C.byref(k), hist.ctypes.data_as(dp), measure, C.byref(out),
                                 nthreads or os.cpu_count() or 1)
    if rc:
        raise RuntimeError("oracle_md_step_nose failed (%d)" % rc)
    return out, k.value
