"""ctypes mirror of include/mdx.h and the loader of the HIP engine library.

The engine library is macsimus_amd/libmdx.so, built in-tree by
macsimus_amd/csrc/Makefile (hipcc, gfx950).  There is no CPU fallback: if
the library is missing, `load_engine()` raises.
"""
import ctypes as C

import numpy as np
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_SO = os.path.join(HERE, "libmdx.so")

MDX_FP64 = 0
MDX_MIXED = 1

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP failure",
          -3: "site out of box", -4: "no convergence", -5: "no GPU device"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)


class LJ(C.Structure):
    """sitesite_t by value (sim/interpot.h:29-57)."""
    _fields_ = [(n, C.c_double) for n in
                ("Sq", "E4", "E24", "E48", "C1q", "C2q", "A", "A4")]


class Erfc(C.Structure):
    """ertab_t (sim/elst.h:62-68), hyperbolic spline interval."""
    _fields_ = [(n, C.c_double) for n in ("Au", "Bu", "Cu", "Ad", "Bd", "Cd")]


class Exp6(C.Structure):
    """the exp-6 form's table entry (sim/exp6/sitesite.h:59-75)."""
    _fields_ = [(n, C.c_double) for n in
                ("tau", "eu", "ef", "ru", "C1q", "C2q", "A", "A4")]


class System(C.Structure):
    _fields_ = [
        ("ns", C.c_int), ("nmol", C.c_int), ("ntypes", C.c_int),
        ("mol_first", _ip), ("site_type", _ip), ("charge", _dp),
        ("excl", _up), ("sstab", C.POINTER(LJ)),
        ("L", C.c_double * 3), ("cutoff", C.c_double),
        ("alpha", C.c_double), ("kappa", C.c_double), ("epsinf", C.c_double),
        ("alphar", C.c_double), ("erfc_sgrid", C.c_double), ("erfc_n", C.c_int),
        ("erfc_tab", C.POINTER(Erfc)),
        ("ncell", C.c_int * 3),
        ("chargepol", _dp), ("alpha_qq", _dp),
        ("mass", _dp), ("ncons", C.c_int), ("cons_first", _ip),
        ("cons_ij", _ip), ("cons_bondq", _dp), ("mol_omega", _dp),
        ("sstab_exp6", C.POINTER(Exp6)),
        ("el_corr", C.c_int), ("el_bg", C.c_int), ("el_diag", C.c_int),
        ("shell_qkappa", _dp), ("shell_rep", _ip),
    ]


class Out(C.Structure):
    _fields_ = [
        ("En_el", C.c_double), ("En_pot", C.c_double), ("En_vir", C.c_double),
        ("Pvir", C.c_double * 6), ("En_kspace", C.c_double),
        ("npair", C.c_longlong), ("nlj", C.c_longlong),
    ]

    def as_dict(self):
        return dict(En_el=self.En_el, En_pot=self.En_pot, En_vir=self.En_vir,
                    Pvir=list(self.Pvir), En_kspace=self.En_kspace,
                    npair=self.npair, nlj=self.nlj)


class Step(C.Structure):
    _fields_ = [
        ("h", C.c_double), ("eps", C.c_double), ("maxit", C.c_int),
        ("thermostat", C.c_int), ("T", C.c_double), ("tau", C.c_double),
        ("nf", C.c_double), ("measure_every", C.c_int),
        ("aspc_k", C.c_int), ("scf_iter", C.c_int), ("scf_omega", C.c_double),
        ("trvp_k", C.c_int), ("nf_tr", C.c_double), ("T_tr_in", C.c_double),
        ("P", C.c_double), ("tau_P", C.c_double), ("Ecorr", C.c_double), ("virial", C.c_int),
    ]


class StepOut(C.Structure):
    _fields_ = [
        ("frc", Out), ("En_self", C.c_double), ("scf_maxerr", C.c_double),
        ("En_kin", C.c_double), ("Pkin", C.c_double * 6),
        ("virc", C.c_double), ("PTvirc", C.c_double * 6),
        ("lambda_", C.c_double), ("shake_iter_max", C.c_int),
        ("En_ext", C.c_double), ("logs", C.c_double), ("dlogs", C.c_double),
        ("En_kin_tr", C.c_double),
        ("L", C.c_double * 3), ("Pref", C.c_double), ("V", C.c_double),
        ("npt_lambda", C.c_double), ("npt_dlambda", C.c_double),
        ("shake_nit", C.c_longlong),
    ]


class Pressure(C.Structure):
    _fields_ = [
        ("vir", C.c_double), ("Pvir", C.c_double * 6), ("Pkin", C.c_double * 6),
        ("Ptens", C.c_double * 6), ("Ptr", C.c_double), ("Ptr_c", C.c_double),
        ("Pevir", C.c_double), ("Pevir_c", C.c_double),
    ]


def declare_setup(lib, prefix):
    """Signatures of the set-up helpers (same in engine and oracle)."""
    f = getattr(lib, prefix + "setss")
    f.argtypes = [C.POINTER(LJ), C.c_double, C.c_double, C.c_double,
                  C.c_double, C.c_int, C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + "setss_exp6")
    f.argtypes = [C.POINTER(Exp6)] + [C.c_double] * 6 + [C.c_int, C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + "setss_buck")
    f.argtypes = [C.POINTER(Exp6)] + [C.c_double] * 6 + [C.c_int, C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + "setss_busing")
    f.argtypes = [C.POINTER(Exp6)] + [C.c_double] * 6 + [C.c_int, C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + "erfc_table")
    f.argtypes = [C.POINTER(Erfc), _ip, _dp, C.c_int, C.c_double, C.c_double,
                  C.c_double, C.c_int]
    f.restype = C.c_int
    f = getattr(lib, prefix + "ewald_autoset")
    f.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double,
                  C.c_double * 3, C.c_double, C.c_double]
    f.restype = C.c_int
    f = getattr(lib, prefix + "ewald_kvectors")
    f.argtypes = [_dp, C.c_double * 3, _ip]
    f.restype = C.c_int
    f = getattr(lib, prefix + "lcsetup")
    f.argtypes = [C.c_int * 3, C.c_int, C.c_double, C.c_double * 3]
    f.restype = C.c_int


_engine = None


def load_engine(path=None):
    """Loads libmdx.so (the HIP engine).  Raises if it is not built."""
    global _engine
    if _engine is not None and path is None:
        return _engine
    so = path or ENGINE_SO
    if not os.path.exists(so):
        raise RuntimeError("MI355X engine library %s is not built "
                           "(run __graft_entry__.build())" % so)
    lib = C.CDLL(so)
    lib.mdx_last_error.restype = C.c_char_p
    lib.mdx_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(System), C.c_int]
    lib.mdx_destroy.argtypes = [C.c_void_p]
    lib.mdx_forces.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(Out), C.c_int]
    lib.mdx_set_box.argtypes = [C.c_void_p, C.c_double * 3]
    lib.mdx_selffield.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(Out), C.c_double, _dp, _dp]
    lib.mdx_pressure.argtypes = [C.POINTER(StepOut), C.c_double * 3, C.c_double, C.c_double,
                                 C.c_int, C.POINTER(Pressure)]
    lib.mdx_cutcorr.argtypes = [C.POINTER(System), _dp, _dp]
    lib.mdx_upload_state.argtypes = [C.c_void_p, _dp, _dp]
    lib.mdx_download_state.argtypes = [C.c_void_p, _dp, _dp]
    lib.mdx_md_steps.argtypes = [C.c_void_p, C.POINTER(Step), C.c_int,
                                 C.POINTER(StepOut)]
    declare_setup(lib, "mdx_")
    from .engine import declare_extras
    declare_extras(lib)
    if path is None:
        _engine = lib
    return lib


def check(lib, rc, what):
    if rc != 0:
        msg = lib.mdx_last_error() if hasattr(lib, "mdx_last_error") else b""
        raise RuntimeError("%s failed: %s (%s)" % (
            what, ERRORS.get(rc, rc), (msg or b"").decode()))


def pressure(lib, step_out, L, Pkinq=1.0, Ecorr=0.0, polar=False):
    """measurePconstraints() of the reference on one measured step's
    outputs (program units)"""
    p = Pressure()
    check(lib, lib.mdx_pressure(C.byref(step_out), (C.c_double * 3)(*L), Pkinq, Ecorr,
                                int(polar), C.byref(p)), "mdx_pressure")
    return p


def cutcorr(lib, built_system):
    """cutcorr(): En.corr of the system from the per-type-pair corrections
    of mdx_setss"""
    corr = np.ascontiguousarray(built_system.info["corr"], dtype=np.float64)
    e = C.c_double(0)
    check(lib, lib.mdx_cutcorr(C.byref(built_system.struct),
                               corr.ctypes.data_as(C.POINTER(C.c_double)), C.byref(e)), "mdx_cutcorr")
    return e.value
