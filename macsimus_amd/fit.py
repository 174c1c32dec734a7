"""Force-field fitting on the engine: force matching (and energy gradients)
for per-type-pair Lennard-Jones sigma/epsilon, or exp-6 sigma/epsilon/a,
and per-type charges.

Parameters theta = [sigma_p (A), epsilon_p (kJ/mol) (, a_p for exp-6) for
each fitted type pair p] + [charge parameters (e)].  A type pair (a,b)
with its own parameters enters cook's site-site table through setss()
with both sites given the same (EvdW, RvdW(, a)), which the combining rule
maps to themselves (sim/lj/sitesite.c:66-112, sim/exp6/sitesite.c:60-76);
for exp-6, sigma is the distance of the minimum.  The type charges are an
affine function of the charge parameters (e.g. water H = -q_O/2).

Gradient of the loss L = sum_frames sum_i |F_i(theta) - F_i^target|^2:
  * LJ / exp-6: the engine's adjoint kernel (mdx_adjoint_lj,
    mdx_adjoint_exp6) gives, for the weights w_i = dL/dF_i = 2 (F_i -
    F_i^target), the derivatives with respect to the table constants of
    each type pair ((Sq, E48, A4, C2q), resp. (tau, ef, ru, A4, C2q)); the
    chain to the parameters goes through setss() (mdx_setss,
    mdx_setss_exp6, which include the smooth-cutoff fixed point) by
    central differences of the constants on the host.
  * charges: the engine's charge adjoint (mdx_adjoint_charges) gives
    sum_i w_i . dF_i/dq_s and dE/dq_s for every site s; the chain to the
    charge parameters is the affine charge map.  The site forces and the
    energy are quadratic forms in the charges (r-space, exclusion, k-space
    and dipole terms alike), so the central difference of the forward
    call, [F(q + h) - F(q - h)] / 2h, is the exact derivative up to
    rounding for any step h: it checks the adjoint, and it is the method
    for systems with Drudes, which the charge adjoint does not take.

The fit (fit_lbfgs) runs L-BFGS in the variables FitVars: C12 and C6 of
each Lennard-Jones pair, on which the forces depend linearly, and the
charge parameters, preconditioned by the Cholesky factor of the
Gauss-Newton matrix J^T J (ForceMatch.preconditioner).

`torch_loss(fm)` wraps the loss as a torch.autograd.Function on CPU
tensors (torch is imported there only: its bundled HIP runtime must not
share a process that computes on the GPU through torch).
"""
import ctypes as C

import numpy as np

from . import abi
from .engine import Engine
from .system import EUNIT_KCAL, ELECTRON, build_system

KCAL_PER_KJ = 1 / 4.184
TWO16 = 2 ** (1 / 6)


class ForceMatch:
    """frames: [nframe][ns][3] positions (for Drude systems [2 ns] rows);
    targets: [nframe][ns][3] target forces on the sites (program units);
    lj_pairs: list of type-name pairs (a, b) whose sigma, epsilon are
    fitted; charge_map: (names of the types, base charges [e] per type,
    matrix [ntypes][nq] so that q_type = base + M @ theta_q)."""

    def __init__(self, ff, counts, L, cutoff, frames, targets, lj_pairs, charge_map,
                 mode=abi.MDX_FP64, epsr=0.05, epsk=0.05, alpha=None, kappa=None):
        form = ff.get("form", "lj")
        if form not in ("lj", "exp6"):
            raise ValueError("ForceMatch fits the Lennard-Jones and exp-6 forms, not %r" % form)
        self.ff = ff
        self.counts = counts
        self.L = [L] * 3 if np.isscalar(L) else list(L)
        self.cutoff = cutoff
        self.frames = [np.ascontiguousarray(f, dtype=np.float64) for f in frames]
        self.targets = [np.ascontiguousarray(t, dtype=np.float64) for t in targets]
        self.bs = build_system(ff, counts, self.L, cutoff, epsr=epsr, epsk=epsk,
                               alpha=alpha, kappa=kappa)
        self.eng = Engine(self.bs, mode)
        self.lib = self.eng.lib
        self.types = list(ff["types"])
        self.nt = len(self.types)
        self.pairs = [(self.types.index(a), self.types.index(b)) for a, b in lj_pairs]
        names, base, M = charge_map
        self.q_types = [self.types.index(n) for n in names]
        self.q_base = np.asarray(base, dtype=np.float64)
        self.q_M = np.asarray(M, dtype=np.float64)
        self.exp6 = form == "exp6"
        self.pp = 3 if self.exp6 else 2          # parameters per fitted pair
        self.npl = self.pp * len(self.pairs)
        self.npar = self.npl + self.q_M.shape[1]
        self.site_type = self.bs.info["site_type"]
        self.polar = bool(self.bs.info["polar"])
        self.qpol = self.bs.info["chargepol"]
        self.ns = self.bs.struct.ns
        self.C2 = cutoff

    # ---- parameters -> engine tables
    def theta0(self):
        """the parameters of the force field the system was built with"""
        th = []
        E = self.ff["type_EvdW_kcal"]
        R = self.ff["type_RvdW"]
        for a, b in self.pairs:
            ea, eb = -E[a] / KCAL_PER_KJ, -E[b] / KCAL_PER_KJ
            if self.exp6:
                # exp-6 rule: sigma_ab = R_a + R_b, eps, a geometric means
                A = self.ff["type_parm"]
                th += [R[a] + R[b], np.sqrt(ea * eb), np.sqrt(A[a] * A[b])]
                continue
            # Lorentz-Berthelot: sigma_ab = (sigma_a + sigma_b)/2, eps = sqrt(eps_a eps_b)
            sa, sb = 2 * R[a] / TWO16, 2 * R[b] / TWO16
            th += [(sa + sb) / 2, np.sqrt(ea * eb)]
        q = self._type_charges_from_ff()
        # least-squares inverse of the affine charge map
        rhs = np.array([q[t] for t in self.q_types]) - self.q_base
        th += list(np.linalg.lstsq(self.q_M, rhs, rcond=None)[0])
        return np.array(th)

    def _type_charges_from_ff(self):
        q = np.zeros(self.nt)
        for sp in self.ff["species"]:
            for t, c in zip(sp["site_type"], sp["site_charge_e"]):
                q[t] = c
        return q

    def _setss(self, sigma, eps_kj):
        e_k = -eps_kj * KCAL_PER_KJ * EUNIT_KCAL
        r = sigma * TWO16 / 2
        ss = abi.LJ()
        corr = C.c_double(0)
        abi.check(self.lib, self.lib.mdx_setss(C.byref(ss), e_k, r, e_k, r, 1, self.C2, C.byref(corr)),
                  "mdx_setss")
        return ss

    def _setss6(self, sigma, eps_kj, a):
        e_k = -eps_kj * KCAL_PER_KJ * EUNIT_KCAL
        ss = abi.Exp6()
        corr = C.c_double(0)
        abi.check(self.lib, self.lib.mdx_setss_exp6(C.byref(ss), e_k, sigma / 2, a, e_k, sigma / 2, a,
                                                    1, self.C2, C.byref(corr)), "mdx_setss_exp6")
        return ss

    def _pair_table(self, th):
        return self._setss6(*th) if self.exp6 else self._setss(*th)

    def _table(self, theta):
        tab = ((abi.Exp6 if self.exp6 else abi.LJ) * (self.nt * self.nt))()
        C.memmove(tab, self.bs.info["sstab"], C.sizeof(tab))
        for k, (a, b) in enumerate(self.pairs):
            ss = self._pair_table(theta[self.pp * k:self.pp * (k + 1)])
            tab[a * self.nt + b] = ss
            tab[b * self.nt + a] = ss
        return tab

    def _charges(self, theta):
        """centre charges [ns] (program units) for the charge parameters;
        for Drude sites the centre carries total - Drude"""
        qt = self._type_charges_from_ff()
        qt[self.q_types] = self.q_base + self.q_M @ theta[self.npl:]
        q = qt[self.site_type] * ELECTRON
        return q - self.qpol if self.polar else q

    def apply(self, theta):
        self.eng.set_params(sstab=self._table(theta), charge=self._charges(theta))

    # ---- loss and gradient
    def forces(self, theta, measure=0):
        self.apply(theta)
        return [self.eng.forces(r, measure=measure) for r in self.frames]

    def loss(self, theta):
        L = 0.0
        for (f, _), t in zip(self.forces(theta), self.targets):
            L += float(((f[:self.ns] - t) ** 2).sum())
        return L

    # table constants of the adjoint: force derivatives, energy derivatives
    # (positions in the vector of _dconst)
    CONSTS = {False: ("Sq", "E48", "A4", "C2q", "E4", "A"),
              True: ("tau", "ef", "ru", "A4", "C2q", "eu", "A")}
    FROWS = {False: [0, 1, 2, 3], True: [0, 1, 2, 3, 4]}
    EROWS = {False: [0, 4, 5, 3], True: [0, 5, 2, 6, 4]}

    def _dconst(self, th):
        """d(table constants)/d(pair parameters) by central differences of
        setss (rows as CONSTS)"""
        names = self.CONSTS[self.exp6]

        def vec(t):
            ss = self._pair_table(t)
            return np.array([getattr(ss, n) for n in names])
        th = np.asarray(th, dtype=np.float64)
        out = np.zeros((len(names), len(th)))
        for k in range(len(th)):
            h = 1e-6 * th[k]
            e = np.zeros(len(th))
            e[k] = h
            out[:, k] = (vec(th + e) - vec(th - e)) / (2 * h)
        return out

    def _adjoint(self, r, w):
        return self.eng.adjoint_exp6(r, w) if self.exp6 else self.eng.adjoint_lj(r, w)

    def _dq_dtheta(self):
        """d(site charge, program units)/d(charge parameter) [ns][nq]"""
        D = np.zeros((self.ns, self.q_M.shape[1]))
        st = np.asarray(self.site_type)[:self.ns]
        for t, ty in enumerate(self.q_types):
            D[st == ty] = self.q_M[t] * ELECTRON
        return D

    def charge_method(self):
        """"adjoint" (mdx_adjoint_charges) unless the system has Drudes,
        where the charge gradient is taken by exact differences"""
        return "differences" if self.polar else "adjoint"

    def loss_and_grad(self, theta, hq=1e-3, charges=None):
        """L(theta) and dL/dtheta (adjoint for LJ; for the charges the
        charge adjoint, or with charges="differences" the central
        differences of the forward forces, exact for these quadratic forms)"""
        theta = np.asarray(theta, dtype=np.float64)
        charges = charges or self.charge_method()
        fr = self.forces(theta)
        L = 0.0
        gF = 0.0
        gq = 0.0
        ws = []
        for (f, _), t, r in zip(fr, self.targets, self.frames):
            d = f[:self.ns] - t
            L += float((d * d).sum())
            w = 2 * d
            ws.append(w)
            gF = gF + self._adjoint(r, w)[0]
            if charges == "adjoint":
                gq = gq + self.eng.adjoint_charges(r, w)[0]
        g = np.zeros(self.npar)
        pp = self.pp
        for k, (a, b) in enumerate(self.pairs):
            gc = gF[a, b] + (gF[b, a] if a != b else 0)          # d L / d(force constants)
            J = self._dconst(theta[pp * k:pp * (k + 1)])[self.FROWS[self.exp6]]
            g[pp * k:pp * (k + 1)] = gc @ J
        if charges == "adjoint":
            g[self.npl:] = gq @ self._dq_dtheta()
        else:
            for m in range(self.q_M.shape[1]):
                e = np.zeros(self.npar)
                e[self.npl + m] = hq
                fp = self.forces(theta + e)
                fm = self.forces(theta - e)
                s = 0.0
                for w, (a_, _), (b_, _) in zip(ws, fp, fm):
                    s += float((w * (a_[:self.ns] - b_[:self.ns])).sum()) / (2 * hq)
                g[self.npl + m] = s
        self.apply(theta)
        return L, g

    def energy_and_grad(self, theta, hq=1e-3, charges=None):
        """sum over frames of En.pot + En.el and its gradient (adjoint gE for
        LJ; for the charges the charge adjoint or exact differences, as
        loss_and_grad)"""
        theta = np.asarray(theta, dtype=np.float64)
        charges = charges or self.charge_method()
        outs = self.forces(theta, measure=1)
        E = sum(o.En_pot + o.En_el for _, o in outs)
        gE = 0.0
        gq = 0.0
        zero = np.zeros((self.ns, 3))
        for r in self.frames:
            gE = gE + self._adjoint(r, zero)[1]
            if charges == "adjoint":
                gq = gq + self.eng.adjoint_charges(r, zero)[1]
        g = np.zeros(self.npar)
        pp = self.pp
        for k, (a, b) in enumerate(self.pairs):
            gc = gE[a, b] + (gE[b, a] if a != b else 0)          # d E / d(energy constants)
            J = self._dconst(theta[pp * k:pp * (k + 1)])
            g[pp * k:pp * (k + 1)] = gc @ J[self.EROWS[self.exp6]]
        if charges == "adjoint":
            g[self.npl:] = gq @ self._dq_dtheta()
        else:
            for m in range(self.q_M.shape[1]):
                e = np.zeros(self.npar)
                e[self.npl + m] = hq
                Ep = sum(o.En_pot + o.En_el for _, o in self.forces(theta + e, measure=1))
                Em = sum(o.En_pot + o.En_el for _, o in self.forces(theta - e, measure=1))
                g[self.npl + m] = (Ep - Em) / (2 * hq)
        self.apply(theta)
        return E, g

    def jtj(self, theta, frames=None, hrel=1e-4):
        """J^T J of the force residuals at theta over the given frame
        indices (default all), J by central differences of the forward
        forces: h = hrel |theta_k| (exact for the charges and epsilon, on
        which the forces depend quadratically resp. linearly)"""
        theta = np.asarray(theta, dtype=np.float64)
        idx = range(len(self.frames)) if frames is None else frames
        J = np.zeros((len(idx), self.ns * 3, self.npar))
        for k in range(self.npar):
            h = hrel * max(abs(theta[k]), 1e-3)
            e = np.zeros(self.npar)
            e[k] = h
            self.apply(theta + e)
            fp = [self.eng.forces(self.frames[i])[0][:self.ns].ravel() for i in idx]
            self.apply(theta - e)
            fm = [self.eng.forces(self.frames[i])[0][:self.ns].ravel() for i in idx]
            for n in range(len(idx)):
                J[n, :, k] = (fp[n] - fm[n]) / (2 * h)
        self.apply(theta)
        return np.einsum("nik,nil->kl", J, J)

    def preconditioner(self, theta, fv, frames=None):
        """R with J_u^T J_u = R^T R (Cholesky), J_u the Jacobian of the force
        residuals over the given frames (default all, scaled to all) with
        respect to the fit variables u of fv (FitVars): in x = R (u - u0)
        the Gauss-Newton Hessian of the loss is 2 I, so every parameter is
        equally well conditioned for a quasi-Newton fit"""
        A = self.jtj(theta, frames)
        A *= len(self.frames) / (len(self.frames) if frames is None else len(frames))
        D = fv.jac(np.asarray(theta, dtype=np.float64))
        return np.linalg.cholesky(D.T @ A @ D).T

    def close(self):
        self.eng.close()


class FitVars:
    """Fit variables u of the parameters theta of a ForceMatch, relative to
    theta0.  A Lennard-Jones pair has u = (C12 / C12_0, C6 / C6_0), C12 =
    4 eps sigma^12, C6 = 4 eps sigma^6: its forces are linear in u (up to
    the smooth-cutoff constants), where in (sigma, eps) the repulsive wall
    goes as sigma^12 and its sigma and eps columns are nearly parallel.  An
    exp-6 pair has u = log sigma, log eps, log a; a charge parameter
    u = theta.  max_step(u, du): the largest t for a trial step u + t du
    that lowers no C12 or C6 below half its value (keeping them positive),
    changes no exp-6 parameter by more than a factor e^0.5 and no charge
    parameter by more than max_q e."""

    def __init__(self, fm, theta0, max_q=0.2):
        self.npl, self.npar, self.exp6 = fm.npl, fm.npar, fm.exp6
        th = np.asarray(theta0, dtype=np.float64)
        s, e = th[0:self.npl:2], th[1:self.npl:2]
        self.c0 = None if self.exp6 else np.stack([4 * e * s ** 12, 4 * e * s ** 6], 1).ravel()
        self.max_q = max_q

    def u(self, theta):
        u = np.array(theta, dtype=np.float64)
        if self.exp6:
            u[:self.npl] = np.log(u[:self.npl])
            return u
        s, e = u[0:self.npl:2].copy(), u[1:self.npl:2].copy()
        u[0:self.npl:2] = 4 * e * s ** 12 / self.c0[0::2]
        u[1:self.npl:2] = 4 * e * s ** 6 / self.c0[1::2]
        return u

    def theta(self, u):
        th = np.array(u, dtype=np.float64)
        if self.exp6:
            th[:self.npl] = np.exp(th[:self.npl])
            return th
        c12, c6 = th[0:self.npl:2] * self.c0[0::2], th[1:self.npl:2] * self.c0[1::2]
        th[0:self.npl:2] = (c12 / c6) ** (1 / 6)
        th[1:self.npl:2] = c6 * c6 / (4 * c12)
        return th

    def jac(self, theta):
        """d theta / d u at theta"""
        D = np.eye(self.npar)
        th = np.asarray(theta, dtype=np.float64)
        if self.exp6:
            D[np.arange(self.npl), np.arange(self.npl)] = th[:self.npl]
            return D
        for k in range(0, self.npl, 2):
            s, e = th[k], th[k + 1]
            c12, c6 = 4 * e * s ** 12, 4 * e * s ** 6
            a12, a6 = self.c0[k], self.c0[k + 1]
            D[k, k], D[k, k + 1] = s / (6 * c12) * a12, -s / (6 * c6) * a6
            D[k + 1, k], D[k + 1, k + 1] = -e / c12 * a12, 2 * e / c6 * a6
        return D

    def max_step(self, u, du):
        t = self.max_q / np.maximum(np.abs(du[self.npl:]), 1e-300)
        if self.exp6:
            t = np.append(t, 0.5 / np.maximum(np.abs(du[:self.npl]), 1e-300))
        else:
            down = du[:self.npl] < 0
            t = np.append(t, 0.5 * u[:self.npl][down] / -du[:self.npl][down])
        return float(t.min()) if len(t) else np.inf


def lbfgs(fun, x0, iters=40, memory=20, maxstep=None, ftol=None, rtol=1e-8, stall=2, log=None):
    """Minimize fun(x) -> (f, g) by L-BFGS: the two-loop recursion for the
    direction (Nocedal & Wright, Numerical Optimization, alg. 7.4, with
    H0 = (s.y / y.y) I) and a backtracking line search on the Armijo
    condition f(x + a d) <= f + 1e-4 a g.d, the first trial a = 1 cut to
    maxstep(x, d) and a point where fun raises RuntimeError (parameters
    the engine refuses) treated as f = inf.  Pairs with s.y <= 1e-12 |s||y|
    are not stored.  Stops after iters iterations, at f <= ftol (if given),
    when no step along d lowers f, or after `stall` iterations in a row
    that lower f by no more than rtol |f| (the rounding floor of f).  Returns (x, [f of every accepted point])."""
    x = np.array(x0, dtype=np.float64)
    f, g = fun(x)
    hist = [f]
    S, Y = [], []
    slow = 0
    for it in range(iters):
        if ftol is not None and f <= ftol:
            break
        q = g.copy()
        al = []
        for s, y in zip(reversed(S), reversed(Y)):
            a = (s @ q) / (y @ s)
            al.append(a)
            q -= a * y
        if S:
            q *= (S[-1] @ Y[-1]) / (Y[-1] @ Y[-1])
        for (s, y), a in zip(zip(S, Y), reversed(al)):
            q += s * (a - (y @ q) / (y @ s))
        d = -q
        gd = g @ d
        if gd >= 0:                       # not a descent direction: restart
            S, Y = [], []
            d, gd = -g, -(g @ g)
        t = 1.0 if maxstep is None else min(1.0, maxstep(x, d))
        for _ in range(30):
            try:
                fn, gn = fun(x + t * d)
            except RuntimeError:
                fn, gn = np.inf, None
            if np.isfinite(fn) and fn <= f + 1e-4 * t * gd:
                break
            # minimum of the quadratic through f, g.d and fn, kept in [0.1 t, 0.5 t]
            tq = -gd * t * t / (2 * (fn - f - gd * t)) if np.isfinite(fn) else 0.1 * t
            t = min(max(tq, 0.1 * t), 0.5 * t)
        else:
            if log:
                log("fit: no descent along the search direction at iteration %d" % it)
            break
        s, y = t * d, gn - g
        if s @ y > 1e-12 * np.linalg.norm(s) * np.linalg.norm(y):
            S.append(s)
            Y.append(y)
            if len(S) > memory:
                S.pop(0)
                Y.pop(0)
        slow = slow + 1 if f - fn <= rtol * abs(f) else 0
        x, f, g = x + s, fn, gn
        hist.append(f)
        if log:
            log("fit: iteration %d, f %.6e, step %.3g, |g| %.3e" % (it + 1, f, t, np.linalg.norm(g)))
        if slow >= stall:
            break
    return x, hist


def fit_lbfgs(fm, theta0, R=None, iters=40, log=None, fv=None, ftol=1e-16, rtol=1e-4):
    """L-BFGS of the force-matching loss of fm in the variables
    x = R (u - u0) / sqrt(L0/2), u the FitVars of theta (default FitVars(fm,
    theta0)), L0 the loss at theta0 and R from fm.preconditioner (computed
    here over all frames if not given): the Gauss-Newton model of L/L0 in
    x has the identity as Hessian, so the first step has the right length
    and every parameter is equally well conditioned.  Trial steps are
    limited by fv.max_step; the fit stops at L <= ftol L0 or after two
    iterations in a row that lower L by no more than rtol L (the
    rounding floor of the engine's loss).  Returns (theta, loss history of the
    accepted points)."""
    theta0 = np.asarray(theta0, dtype=np.float64)
    fv = FitVars(fm, theta0) if fv is None else fv
    if R is None:
        R = fm.preconditioner(theta0, fv)
    u0 = fv.u(theta0)
    L0, _ = fm.loss_and_grad(theta0)
    Rinv = np.linalg.inv(R) * np.sqrt(L0 / 2)

    def fun(x):
        th = fv.theta(u0 + Rinv @ x)
        L, g = fm.loss_and_grad(th)
        return L / L0, Rinv.T @ (fv.jac(th).T @ g) / L0

    def maxstep(x, d):
        return fv.max_step(u0 + Rinv @ x, Rinv @ d)
    x, hist = lbfgs(fun, np.zeros(fm.npar), iters=iters, maxstep=maxstep, ftol=ftol, rtol=rtol,
                    log=log)
    return fv.theta(u0 + Rinv @ x), [h * L0 for h in hist]


def torch_loss(loss_and_grad):
    """A torch function theta -> L(theta) whose backward returns the given
    gradient: loss_and_grad(theta_numpy) -> (L, dL/dtheta).  CPU tensors."""
    import torch

    class _ForceMatchLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            L, g = loss_and_grad(theta.detach().cpu().numpy().astype(np.float64))
            ctx.save_for_backward(torch.as_tensor(np.asarray(g), dtype=theta.dtype))
            return theta.new_tensor(L)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g

    return _ForceMatchLoss.apply


# ---- the force-matching workload: 64 frames of a 23,712-site box
FIT_PAIRS = [("O", "O"), ("NA", "O"), ("CL", "O"), ("NA", "CL")]
# q_H = -q_O/2 (neutral water), q_O, q_Na, q_Cl
FIT_CHARGES = (["H", "O", "NA", "CL"], [0, 0, 0, 0],
               [[-0.5, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])


def make_frames(nframes=64, equil=1000, every=20, seed_shift=0):
    """Frames of the 23,712-site SPC/E + 0.5 mol/kg NaCl box, equilibrated
    by the engine (FP64, Berendsen 298 K, h = 2 fs, SHAKE) from the lattice:
    equil steps, then one frame every `every` steps."""
    from . import workloads
    wl = workloads.brine_c5()
    bs = build_system(wl["ff"], wl["counts"], [wl["L"]] * 3, wl["cutoff"], epsr=0.05, epsk=0.05)
    eng = Engine(bs, abi.MDX_FP64)
    eng.autotune(wl["r"], reps=1)
    nf = workloads.degrees_of_freedom(wl["ff"], wl["counts"])
    p = abi.Step(h=0.002, eps=1e-7, maxit=0, thermostat=1, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=0)
    eng.upload_state(wl["r"], np.zeros_like(wl["r"]))
    eng.md_steps(p, equil)
    frames = []
    for _ in range(nframes):
        eng.md_steps(p, every)
        frames.append(eng.download_state()[0])
    info = dict(alpha=bs.info["alpha"], kappa=bs.info["kappa"])
    eng.close()
    return wl, frames, info


def main():
    import argparse
    import json
    import sys
    import time

    ap = argparse.ArgumentParser(description="force matching of LJ sigma/epsilon and charges")
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--iters", type=int, default=40, help="L-BFGS iterations")
    ap.add_argument("--no-fd", action="store_true", help="skip the finite-difference check")
    a = ap.parse_args()
    t0 = time.perf_counter()
    wl, frames, info = make_frames(a.frames)
    t_frames = time.perf_counter() - t0
    kw = dict(alpha=info["alpha"], kappa=info["kappa"])
    ref = ForceMatch(wl["ff"], wl["counts"], wl["L"], wl["cutoff"], frames,
                     [np.zeros_like(f) for f in frames], FIT_PAIRS, FIT_CHARGES, **kw)
    th0 = ref.theta0()
    # targets: +5 % sigma and epsilon, charges -0.02 e (O), +0.02 e (Na), -0.02 e (Cl)
    tht = th0 * np.array([1.05] * 8 + [1, 1, 1]) + np.array([0] * 8 + [-0.02, 0.02, -0.02])
    targets = [f for f, _ in ref.forces(tht)]
    ref.close()
    res = {"workload": wl["name"], "frames": a.frames, "sites": int(sum(
        n * len(sp["site_type"]) for n, sp in zip(wl["counts"], wl["ff"]["species"]))),
        "frames_s": t_frames, "theta0": th0.tolist(), "theta_target": tht.tolist()}
    fms = {m: ForceMatch(wl["ff"], wl["counts"], wl["L"], wl["cutoff"], frames, targets,
                         FIT_PAIRS, FIT_CHARGES, mode=m, **kw) for m in (abi.MDX_FP64, abi.MDX_MIXED)}
    fm64, fm32 = fms[abi.MDX_FP64], fms[abi.MDX_MIXED]
    t0 = time.perf_counter()
    L64, g64 = fm64.loss_and_grad(th0)
    res["ms_loss_grad_fp64"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    L32, g32 = fm32.loss_and_grad(th0)
    res["ms_loss_grad_mixed"] = (time.perf_counter() - t0) * 1e3
    res["loss0"] = L64
    res["grad_mixed_vs_fp64"] = float(np.abs(g32 - g64).max() / np.abs(g64).max())
    if not a.no_fd:
        gfd = np.zeros_like(g64)
        for k in range(len(th0)):
            h = 1e-4 * abs(th0[k])
            e = np.zeros_like(th0)
            e[k] = h
            gfd[k] = (fm64.loss(th0 + e) - fm64.loss(th0 - e)) / (2 * h)
        res["grad_vs_central_fd"] = float(np.abs(g64 - gfd).max() / np.abs(gfd).max())
    print("gradients done: %s" % json.dumps({k: v for k, v in res.items() if k.startswith(("grad", "ms", "loss"))}),
          file=sys.stderr, flush=True)

    # L-BFGS on the mixed engine in preconditioned C12/C6 and charge
    # variables (Gauss-Newton J^T J of 4 frames at theta0)
    t0 = time.perf_counter()
    fv = FitVars(fm32, th0)
    R = fm32.preconditioner(th0, fv, frames=list(range(min(4, a.frames))))
    res["precond_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    th, hist = fit_lbfgs(fm32, th0, R, iters=a.iters, fv=fv,
                         log=lambda m: print(m, file=sys.stderr, flush=True))
    res["fit_s"] = time.perf_counter() - t0
    res["loss_history_rel"] = [h / L64 for h in hist]
    res["theta_fit"] = th.tolist()
    res["fit_param_rel_err"] = (np.abs(th - tht) / np.abs(tht)).tolist()
    res["fit_param_rel_err_max"] = float((np.abs(th - tht) / np.abs(tht)).max())
    for fm in fms.values():
        fm.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
