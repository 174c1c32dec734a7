"""Python handle on the MI355X engine (libmdx.so) through its C ABI.

Mirrors cook's force call: `Engine.forces(r, measure)` is forces(B,A) of
cook/forces.c:93-224 with A->rp = r and B->rp = the returned forces.  For
a Drude (POLAR) system the configuration has 2*ns rows: the centres, then
the Drude offsets (sim/simglob.h polarrof()).
"""
import ctypes as C

import numpy as np

from . import abi


class Engine:
    def __init__(self, built_system, mode=abi.MDX_MIXED, lib=None):
        self.lib = lib or abi.load_engine()
        self.bs = built_system            # keeps the arrays alive
        self.ns = built_system.struct.ns
        self.polar = bool(built_system.info.get("polar"))
        self.nrow = 2 * self.ns if self.polar else self.ns
        self.ctx = C.c_void_p()
        rc = self.lib.mdx_create(C.byref(self.ctx), C.byref(built_system.struct), mode)
        abi.check(self.lib, rc, "mdx_create")
        self.mode = mode
        self._rdf_grid = 0

    def resize(self, built_system):
        """another system in this context (mdx_resize): the settings carry
        over, the MD state must be uploaded again"""
        rc = self.lib.mdx_resize(self.ctx, C.byref(built_system.struct))
        abi.check(self.lib, rc, "mdx_resize")
        self.bs = built_system
        self.ns = built_system.struct.ns
        self.polar = bool(built_system.info.get("polar"))
        self.nrow = 2 * self.ns if self.polar else self.ns

    def close(self):
        if self.ctx:
            self.lib.mdx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forces(self, r, measure=1):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.nrow, 3)
        f = np.empty_like(r)
        out = abi.Out()
        rc = self.lib.mdx_forces(self.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                 f.ctypes.data_as(C.POINTER(C.c_double)), C.byref(out),
                                 int(measure))
        abi.check(self.lib, rc, "mdx_forces")
        return f, out

    def selffield(self, r, f, out, omega):
        """one selffield() iteration in place on r, f (2*ns rows); returns
        (En.self, max SCF error)"""
        assert r.flags.c_contiguous and f.flags.c_contiguous and r.shape == (self.nrow, 3)
        es = C.c_double(0)
        me = C.c_double(0)
        rc = self.lib.mdx_selffield(self.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                    f.ctypes.data_as(C.POINTER(C.c_double)), C.byref(out),
                                    float(omega), C.byref(es), C.byref(me))
        abi.check(self.lib, rc, "mdx_selffield")
        return es.value, me.value

    def set_aspc(self, k, kk=0, hist=None):
        ptr = None
        if hist is not None:
            hist = np.ascontiguousarray(hist, dtype=np.float64)
            assert hist.shape == (k, self.ns, 3)
            ptr = hist.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.mdx_set_aspc(self.ctx, int(k), int(kk), ptr)
        abi.check(self.lib, rc, "mdx_set_aspc")

    def get_aspc(self):
        k = C.c_int(0)
        kk = C.c_int(0)
        rc = self.lib.mdx_get_aspc(self.ctx, C.byref(k), C.byref(kk), None)
        abi.check(self.lib, rc, "mdx_get_aspc")
        hist = np.zeros((max(k.value, 0), self.ns, 3))
        if k.value > 0:
            rc = self.lib.mdx_get_aspc(self.ctx, C.byref(k), C.byref(kk),
                                       hist.ctypes.data_as(C.POINTER(C.c_double)))
            abi.check(self.lib, rc, "mdx_get_aspc")
        return k.value, kk.value, hist

    def set_nose(self, logs=0.0, logs1=0.0, k=0, kk=0, hist=None):
        """Nose-Hoover state: logs, logs1 = logs(t)-logs(t-h), predictor
        length k, length reached kk, history [k][1+3ns] (None: zeros)"""
        ptr = None
        if hist is not None:
            hist = np.ascontiguousarray(hist, dtype=np.float64)
            assert hist.shape == (k, 1 + 3 * self.ns)
            ptr = hist.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.mdx_set_nose(self.ctx, float(logs), float(logs1), int(k), int(kk), ptr)
        abi.check(self.lib, rc, "mdx_set_nose")

    def get_nose(self):
        """(logs, logs1, k, kk, hist [k][1+3ns])"""
        a, b, k, kk = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        rc = self.lib.mdx_get_nose(self.ctx, C.byref(a), C.byref(b), C.byref(k), C.byref(kk), None)
        abi.check(self.lib, rc, "mdx_get_nose")
        hist = np.zeros((k.value, 1 + 3 * self.ns))
        if k.value:
            rc = self.lib.mdx_get_nose(self.ctx, None, None, None, None,
                                       hist.ctypes.data_as(C.POINTER(C.c_double)))
            abi.check(self.lib, rc, "mdx_get_nose")
        return a.value, b.value, k.value, kk.value, hist

    def set_npt(self, logs=0.0, logs1=0.0, lambda1=0.0, k=0, kk=0, hist=None):
        """isotropic NPT state: logs, logs1, lambda1 = log L(t) - log L(t-h),
        predictor length k, length reached kk, history [k][2+3ns] (per
        level logs1, lambda1, the sites' r1; None: zeros)"""
        ptr = None
        if hist is not None:
            hist = np.ascontiguousarray(hist, dtype=np.float64)
            assert hist.shape == (k, 2 + 3 * self.ns)
            ptr = hist.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.mdx_set_npt(self.ctx, float(logs), float(logs1), float(lambda1), int(k),
                                  int(kk), ptr)
        abi.check(self.lib, rc, "mdx_set_npt")

    def get_npt(self):
        """(logs, logs1, lambda1, k, kk, hist [k][2+3ns])"""
        a, b, l1, k, kk = C.c_double(), C.c_double(), C.c_double(), C.c_int(), C.c_int()
        rc = self.lib.mdx_get_npt(self.ctx, C.byref(a), C.byref(b), C.byref(l1), C.byref(k),
                                  C.byref(kk), None)
        abi.check(self.lib, rc, "mdx_get_npt")
        hist = np.zeros((k.value, 2 + 3 * self.ns))
        if k.value:
            rc = self.lib.mdx_get_npt(self.ctx, None, None, None, None, None,
                                      hist.ctypes.data_as(C.POINTER(C.c_double)))
            abi.check(self.lib, rc, "mdx_get_npt")
        return a.value, b.value, l1.value, k.value, kk.value, hist

    def set_box(self, L):
        """a new box (k-vector set kept, ekk recomputed)"""
        rc = self.lib.mdx_set_box(self.ctx, (C.c_double * 3)(*[float(x) for x in L]))
        abi.check(self.lib, rc, "mdx_set_box")

    def box(self):
        """the context's box L (changed by NPT steps and set_box)"""
        L = (C.c_double * 3)()
        rc = self.lib.mdx_get_box(self.ctx, L)
        abi.check(self.lib, rc, "mdx_get_box")
        return np.array(L[:])

    def cook_cells(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        cid = np.empty(self.ns, np.int32)
        rc = self.lib.mdx_cook_cells(self.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                     cid.ctypes.data_as(C.POINTER(C.c_int)))
        abi.check(self.lib, rc, "mdx_cook_cells")
        return cid

    def info(self):
        iv = (C.c_int * 20)()
        self.lib.mdx_info(self.ctx, iv, 20)
        keys = ("ncx", "ncy", "ncz", "stencil", "nq", "krows", "kblocks", "ksegments",
                "ktriples", "ksplits", "kxmax", "mode", "kchunk", "nkchunks", "kparts_asked",
                "kparts", "pair_cols", "kmfma", "kpass1_tiles", "kprod")
        return dict(zip(keys, list(iv)))

    def set_kchunk(self, kzc, parts=16):
        rc = self.lib.mdx_set_kchunk(self.ctx, int(kzc), int(parts))
        abi.check(self.lib, rc, "mdx_set_kchunk")

    def autotune(self, r=None, reps=3):
        """picks the fastest cell stencil and k-space kz chunk on r (None:
        the uploaded MD state); returns (w, kz chunk, k parts, pair columns
        per part, ms pair, ms pass 2)"""
        choice = (C.c_int * 4)()
        ms = (C.c_double * 2)()
        ptr = None
        if r is not None:
            r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.nrow, 3)
            ptr = r.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.mdx_autotune(self.ctx, ptr, int(reps), choice, ms)
        abi.check(self.lib, rc, "mdx_autotune")
        return choice[0], choice[1], choice[2], choice[3], ms[0], ms[1]

    def set_timing(self, on=True):
        self.lib.mdx_set_timing(self.ctx, int(on))

    FAMILIES = ("bin", "pair", "excl", "kpass1", "kcoef", "kpass2", "shake")

    def timing(self):
        """{family: (total ms, launches)} since the last call (HIP events)."""
        ms = (C.c_double * 7)()
        cnt = (C.c_int * 7)()
        rc = self.lib.mdx_get_timing(self.ctx, ms, cnt, 7)
        abi.check(self.lib, rc, "mdx_get_timing")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.FAMILIES)}

    def set_params(self, sstab=None, charge=None, chargepol=None):
        """new site-site table (ctypes array of abi.LJ, or of abi.Exp6 for
        an exp-6 context, ntypes^2) and/or charges"""
        dp = C.POINTER(C.c_double)
        if sstab is not None and getattr(sstab, "_type_", None) is abi.Exp6:
            rc = self.lib.mdx_set_params_exp6(self.ctx, C.cast(sstab, C.POINTER(abi.Exp6)))
            abi.check(self.lib, rc, "mdx_set_params_exp6")
            sstab = None
            if charge is None and chargepol is None:
                return
        q = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        qp = None if chargepol is None else np.ascontiguousarray(chargepol, dtype=np.float64)
        rc = self.lib.mdx_set_params(self.ctx, None if sstab is None else C.cast(sstab, C.POINTER(abi.LJ)),
                                     None if q is None else q.ctypes.data_as(dp),
                                     None if qp is None else qp.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_set_params")

    def adjoint_exp6(self, r, w):
        """(gF, gE) [nt, nt, 5]: sum_i w_i . dF_i/dc for c = (tau, ef, ru,
        A4, C2q) and dE/dc for c = (tau, eu, ru, A, C2q) of each type pair
        of an exp-6 context (see include/mdx.h)"""
        nt = self.bs.struct.ntypes
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.nrow, 3)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.ns, 3)
        gF = np.zeros((nt, nt, 5))
        gE = np.zeros((nt, nt, 5))
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_adjoint_exp6(self.ctx, r.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                       gF.ctypes.data_as(dp), gE.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_adjoint_exp6")
        return gF, gE

    def adjoint_lj(self, r, w):
        """(gF, gE) [nt, nt, 4]: sum_i w_i . dF_i/dc and dE/dc for the LJ
        constants of each type pair (see include/mdx.h)"""
        nt = self.bs.struct.ntypes
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.nrow, 3)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.ns, 3)
        gF = np.zeros((nt, nt, 4))
        gE = np.zeros((nt, nt, 4))
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_adjoint_lj(self.ctx, r.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                     gF.ctypes.data_as(dp), gE.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_adjoint_lj")
        return gF, gE

    def adjoint_lj_noalloc(self, r, w):
        """adjoint_lj through mdx_adjoint_lj_noalloc: the same results with
        the scratch in the library's static device array (no device
        allocation per call; systems beyond MDX_ADJ_SCRATCH_BYTES raise)"""
        nt = self.bs.struct.ntypes
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.nrow, 3)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.ns, 3)
        gF = np.zeros((nt, nt, 4))
        gE = np.zeros((nt, nt, 4))
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_adjoint_lj_noalloc(self.ctx, r.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                             gF.ctypes.data_as(dp), gE.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_adjoint_lj_noalloc")
        return gF, gE

    def adjoint_charges(self, r, w):
        """(gF, gE) [ns]: sum_k w_k . dF_k/dq_i and dEn_el/dq_i for every
        site charge q_i (program units; see include/mdx.h)"""
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.nrow, 3)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.ns, 3)
        gF = np.zeros(self.ns)
        gE = np.zeros(self.ns)
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_adjoint_charges(self.ctx, r.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                          gF.ctypes.data_as(dp), gE.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_adjoint_charges")
        return gF, gE

    def set_graph(self, on=True):
        rc = self.lib.mdx_set_graph(self.ctx, int(on))
        abi.check(self.lib, rc, "mdx_set_graph")

    def set_kspace_mfma(self, on=True):
        rc = self.lib.mdx_set_kspace_mfma(self.ctx, int(on))
        abi.check(self.lib, rc, "mdx_set_kspace_mfma")

    def set_kspace_products(self, np_=0):
        """pass 2's products: 0 FP32 operands, 3 or 6 split-FP16 products"""
        rc = self.lib.mdx_set_kspace_products(self.ctx, int(np_))
        abi.check(self.lib, rc, "mdx_set_kspace_products")

    def set_shake_species(self, mol_species=None, nspec=1):
        """species index of each molecule for the SHAKE sweep counts and
        omega (None: one species)"""
        if mol_species is None:
            rc = self.lib.mdx_set_shake_species(self.ctx, None, 1)
        else:
            a = np.ascontiguousarray(mol_species, dtype=np.int32)
            self._shake_sp = a
            rc = self.lib.mdx_set_shake_species(self.ctx, a.ctypes.data_as(C.POINTER(C.c_int)),
                                                int(nspec))
        abi.check(self.lib, rc, "mdx_set_shake_species")
        self._shake_nsp = 1 if mol_species is None else int(nspec)

    def set_shake_omega(self, omega):
        """SHAKE relaxation omega (= omegac/2) per species"""
        a = np.ascontiguousarray(np.atleast_1d(omega), dtype=np.float64)
        rc = self.lib.mdx_set_shake_omega(self.ctx, a.ctypes.data_as(C.POINTER(C.c_double)))
        abi.check(self.lib, rc, "mdx_set_shake_omega")

    def shake_nit(self):
        """SHAKE sweeps per species over the last md_steps call"""
        n = getattr(self, "_shake_nsp", 1)
        v = (C.c_longlong * n)()
        rc = self.lib.mdx_get_shake_nit(self.ctx, v, n)
        abi.check(self.lib, rc, "mdx_get_shake_nit")
        return np.array(list(v), dtype=np.int64)

    def set_kspace_pass1(self, tiles=0):
        rc = self.lib.mdx_set_kspace_pass1(self.ctx, int(tiles))
        abi.check(self.lib, rc, "mdx_set_kspace_pass1")

    def set_overlap(self, on=True):
        rc = self.lib.mdx_set_overlap(self.ctx, int(on))
        abi.check(self.lib, rc, "mdx_set_overlap")

    def set_stencil(self, w, pcols=0):
        rc = self.lib.mdx_set_stencil(self.ctx, int(w), int(pcols))
        abi.check(self.lib, rc, "mdx_set_stencil")

    # ---- RDF histograms of the measured calls
    def set_rdf(self, grid, cutoff):
        """start the histograms from zero (cook's rdf.grid, truncated to an
        integer as initrdf() does, and rdf.cutoff); grid 0 turns them off"""
        rc = self.lib.mdx_set_rdf(self.ctx, int(grid), float(cutoff))
        abi.check(self.lib, rc, "mdx_set_rdf")
        self._rdf_grid = grid

    def rdf(self):
        """dict(hist [records, bins] uint64, nmeas, V, grid): records of
        the site-type pairs (i, j<=i) in cook's order, or the one merged
        record for grid < 0"""
        nrec, nh, nm, V = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        rc = self.lib.mdx_get_rdf(self.ctx, None, 0, C.byref(nrec), C.byref(nh), None, None, None)
        abi.check(self.lib, rc, "mdx_get_rdf")
        hist = np.zeros((nrec.value, nh.value), np.uint64)
        rc = self.lib.mdx_get_rdf(self.ctx, hist.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                                  hist.size, None, None, C.byref(nm), C.byref(V))
        abi.check(self.lib, rc, "mdx_get_rdf")
        return dict(hist=hist, nmeas=nm.value, V=V.value, grid=self._rdf_grid)

    # ---- device-resident MD
    def upload_state(self, r, r1):
        r = np.ascontiguousarray(r, dtype=np.float64)
        r1 = np.ascontiguousarray(r1, dtype=np.float64)
        assert r.size == 3 * self.nrow and r1.size == 3 * self.nrow
        rc = self.lib.mdx_upload_state(self.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                       r1.ctypes.data_as(C.POINTER(C.c_double)))
        abi.check(self.lib, rc, "mdx_upload_state")

    def download_state(self):
        r = np.empty((self.nrow, 3))
        r1 = np.empty((self.nrow, 3))
        rc = self.lib.mdx_download_state(self.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                         r1.ctypes.data_as(C.POINTER(C.c_double)))
        abi.check(self.lib, rc, "mdx_download_state")
        return r, r1

    def md_steps(self, params, nsteps):
        out = abi.StepOut()
        rc = self.lib.mdx_md_steps(self.ctx, C.byref(params), int(nsteps), C.byref(out))
        abi.check(self.lib, rc, "mdx_md_steps")
        return out


class Batch:
    """Batched replicas (mdx_batch_*): independent atomic systems of uneven
    size run by one launch per kernel.  Site arrays are lists of [ns_s, 3]
    arrays or their concatenation."""

    def __init__(self, built_systems, mode=abi.MDX_MIXED, lib=None):
        self.lib = lib or abi.load_engine()
        self.bss = list(built_systems)
        self.nsys = len(self.bss)
        self.ns = [b.struct.ns for b in self.bss]
        self.off = np.concatenate([[0], np.cumsum(self.ns)]).astype(int)
        arr = (abi.System * self.nsys)(*[b.struct for b in self.bss])
        self._arr = arr
        self.h = C.c_void_p()
        rc = self.lib.mdx_batch_create(C.byref(self.h), self.nsys, arr, mode)
        abi.check(self.lib, rc, "mdx_batch_create")
        self.mode = mode

    def close(self):
        if self.h:
            self.lib.mdx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _cat(self, rs):
        if isinstance(rs, (list, tuple)):
            rs = np.concatenate([np.asarray(x, np.float64).reshape(-1, 3) for x in rs])
        rs = np.ascontiguousarray(rs, dtype=np.float64).reshape(-1, 3)
        assert rs.shape[0] == self.off[-1]
        return rs

    def split(self, a):
        return [a[self.off[s]:self.off[s + 1]] for s in range(self.nsys)]

    def forces(self, rs, measure=1):
        """per-system forces (list) and mdx_out_t array"""
        r = self._cat(rs)
        f = np.empty_like(r)
        out = (abi.Out * self.nsys)()
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_batch_forces(self.h, r.ctypes.data_as(dp), f.ctypes.data_as(dp), out,
                                       int(measure))
        abi.check(self.lib, rc, "mdx_batch_forces")
        return self.split(f), list(out)

    def upload_state(self, rs, r1s):
        r, r1 = self._cat(rs), self._cat(r1s)
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_batch_upload_state(self.h, r.ctypes.data_as(dp), r1.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_batch_upload_state")

    def download_state(self):
        r = np.empty((self.off[-1], 3))
        r1 = np.empty_like(r)
        dp = C.POINTER(C.c_double)
        rc = self.lib.mdx_batch_download_state(self.h, r.ctypes.data_as(dp), r1.ctypes.data_as(dp))
        abi.check(self.lib, rc, "mdx_batch_download_state")
        return self.split(r), self.split(r1)

    def md_steps(self, params, nsteps):
        """params: list of abi.Step (one per system) or one for all"""
        if isinstance(params, abi.Step):
            params = [params] * self.nsys
        p = (abi.Step * self.nsys)(*params)
        out = (abi.StepOut * self.nsys)()
        rc = self.lib.mdx_batch_md_steps(self.h, p, int(nsteps), out)
        abi.check(self.lib, rc, "mdx_batch_md_steps")
        return list(out)

    def set_graph(self, on=True):
        abi.check(self.lib, self.lib.mdx_batch_set_graph(self.h, int(on)), "mdx_batch_set_graph")

    def set_stencil(self, w):
        abi.check(self.lib, self.lib.mdx_batch_set_stencil(self.h, int(w)), "mdx_batch_set_stencil")

    def set_kspace(self, kzc, kparts):
        abi.check(self.lib, self.lib.mdx_batch_set_kspace(self.h, int(kzc), int(kparts)),
                  "mdx_batch_set_kspace")

    def autotune(self, rs):
        """fastest stencil, k-space kz chunk and parts on rs: (w, kzc,
        kparts, ms after the stencil, ms after the k-space)"""
        r = self._cat(rs)
        ch = (C.c_int * 3)()
        ms = (C.c_double * 2)()
        rc = self.lib.mdx_batch_autotune(self.h, r.ctypes.data_as(C.POINTER(C.c_double)), ch, ms)
        abi.check(self.lib, rc, "mdx_batch_autotune")
        return ch[0], ch[1], ch[2], ms[0], ms[1]

    def set_timing(self, on=True):
        abi.check(self.lib, self.lib.mdx_batch_set_timing(self.h, int(on)), "mdx_batch_set_timing")

    def timing(self):
        """{family: (total ms, launches)} since the last call (HIP events)"""
        ms = (C.c_double * 7)()
        cnt = (C.c_int * 7)()
        rc = self.lib.mdx_batch_get_timing(self.h, ms, cnt, 7)
        abi.check(self.lib, rc, "mdx_batch_get_timing")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(Engine.FAMILIES)}

    def set_kspace_mfma(self, on=True):
        abi.check(self.lib, self.lib.mdx_batch_set_kspace_mfma(self.h, int(on)),
                  "mdx_batch_set_kspace_mfma")

    def info(self):
        iv = (C.c_int * 14)()
        self.lib.mdx_batch_info(self.h, iv, 14)
        return dict(zip(("nsys", "sites", "padded", "cells", "charged", "mode", "stencil",
                         "charges", "kseg", "units1", "units2", "kchunk", "kparts", "kmfma"),
                        list(iv)))


def declare_extras(lib):
    dp = C.POINTER(C.c_double)
    lib.mdx_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(abi.System), C.c_int]
    lib.mdx_batch_destroy.argtypes = [C.c_void_p]
    lib.mdx_batch_forces.argtypes = [C.c_void_p, dp, dp, C.POINTER(abi.Out), C.c_int]
    lib.mdx_batch_upload_state.argtypes = [C.c_void_p, dp, dp]
    lib.mdx_batch_download_state.argtypes = [C.c_void_p, dp, dp]
    lib.mdx_batch_md_steps.argtypes = [C.c_void_p, C.POINTER(abi.Step), C.c_int,
                                       C.POINTER(abi.StepOut)]
    lib.mdx_batch_set_graph.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_batch_set_stencil.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_batch_autotune.argtypes = [C.c_void_p, dp, C.POINTER(C.c_int), dp]
    lib.mdx_batch_set_kspace.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mdx_batch_set_kspace_mfma.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
    lib.mdx_batch_set_timing.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_batch_get_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.c_int]
    lib.mdx_cook_cells.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.mdx_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
    lib.mdx_set_timing.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_get_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]
    lib.mdx_set_stencil.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mdx_set_kchunk.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mdx_set_graph.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_set_overlap.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_set_kspace_mfma.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_set_kspace_pass1.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_set_kspace_products.argtypes = [C.c_void_p, C.c_int]
    lib.mdx_set_shake_species.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
    lib.mdx_set_shake_omega.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    lib.mdx_get_shake_nit.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.c_int]
    lib.mdx_set_params.argtypes = [C.c_void_p, C.POINTER(abi.LJ), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    lib.mdx_adjoint_lj.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
    lib.mdx_adjoint_lj_noalloc.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
    lib.mdx_set_params_exp6.argtypes = [C.c_void_p, C.POINTER(abi.Exp6)]
    lib.mdx_adjoint_exp6.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
    lib.mdx_adjoint_charges.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
    lib.mdx_autotune.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int,
                                 C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.mdx_set_aspc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.mdx_get_aspc.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_double)]
    lib.mdx_set_nose.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.POINTER(C.c_double)]
    lib.mdx_get_nose.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.mdx_set_npt.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                C.POINTER(C.c_double)]
    lib.mdx_get_npt.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]
    lib.mdx_get_box.argtypes = [C.c_void_p, C.c_double * 3]
    lib.mdx_resize.argtypes = [C.c_void_p, C.POINTER(abi.System)]
    lib.mdx_set_rdf.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.mdx_get_rdf.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), C.c_long,
                                C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]
