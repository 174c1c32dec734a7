"""The fit driver of macsimus_amd.fit on the CPU, without the engine: the
fit variables (C12/C6 for Lennard-Jones pairs, logs for exp-6 pairs) against
their closed forms and central differences, the L-BFGS routine on a
quadratic with a known minimum, and the preconditioned fit on a toy force
matching with the same shape as the C5 fit (4 LJ type pairs, pair forces
with Coulomb terms, targets at +5 % sigma and epsilon and +-0.02 e): it
must end at the target parameters."""
import numpy as np
import pytest

from macsimus_amd.fit import FitVars, fit_lbfgs, lbfgs


class Shape:
    def __init__(self, npl, nq, exp6=False):
        self.npl, self.npar, self.exp6 = npl, npl + nq, exp6


SIG = np.array([3.166, 2.758, 3.783, 3.375])
EPS = np.array([0.65, 0.5946, 0.5219, 0.4774])
TH0 = np.concatenate([np.stack([SIG, EPS], 1).ravel(), [-0.8476, 1.0, -1.0]])
THT = TH0 * np.array([1.05] * 8 + [1, 1, 1]) + np.array([0] * 8 + [-0.02, 0.02, -0.02])


def test_fitvars_lj_closed_form_and_jacobian():
    fv = FitVars(Shape(8, 3), TH0)
    th = TH0 * 1.03
    u = fv.u(th)
    # C12 = 4 eps sigma^12 relative to theta0's, C6 likewise; charges as they are
    s, e = th[0:8:2], th[1:8:2]
    np.testing.assert_allclose(u[0:8:2], (e * s ** 12) / (EPS * SIG ** 12), rtol=1e-14)
    np.testing.assert_allclose(u[1:8:2], (e * s ** 6) / (EPS * SIG ** 6), rtol=1e-14)
    np.testing.assert_array_equal(u[8:], th[8:])
    np.testing.assert_allclose(fv.u(TH0)[:8], 1.0, rtol=1e-14)
    np.testing.assert_allclose(fv.theta(u), th, rtol=1e-14)
    D = fv.jac(th)
    Dfd = np.empty_like(D)
    for k in range(len(u)):
        du = np.zeros(len(u))
        du[k] = 1e-6
        Dfd[:, k] = (fv.theta(u + du) - fv.theta(u - du)) / 2e-6
    np.testing.assert_allclose(D, Dfd, atol=1e-8)


def test_fitvars_exp6_logs():
    th0 = np.array([3.55, 0.6, 13.0, 3.9, 0.4, 14.0, 0.8])
    fv = FitVars(Shape(6, 1, exp6=True), th0)
    u = fv.u(th0)
    np.testing.assert_allclose(u[:6], np.log(th0[:6]), rtol=1e-15)
    np.testing.assert_allclose(fv.theta(u), th0, rtol=1e-15)
    np.testing.assert_allclose(np.diag(fv.jac(th0)), np.append(th0[:6], 1.0), rtol=1e-15)
    # a step limited to a factor e^0.5 on the log parameters and 0.2 e on the charge
    assert fv.max_step(u, np.array([1.0, 0, 0, 0, 0, 0, 0])) == pytest.approx(0.5)
    assert fv.max_step(u, np.array([0, 0, 0, 0, 0, 0, -2.0])) == pytest.approx(0.1)


def test_fitvars_max_step_keeps_c12_c6_positive():
    fv = FitVars(Shape(2, 1), TH0[[0, 1, 8]])
    u = np.array([1.0, 2.0, 0.0])
    # lowering C12 is limited to half its value; raising it is not limited
    assert fv.max_step(u, np.array([-1.0, 0, 0])) == pytest.approx(0.5)
    assert fv.max_step(u, np.array([0, -1.0, 0])) == pytest.approx(1.0)
    assert fv.max_step(u, np.array([10.0, 10.0, 0.01])) == pytest.approx(20.0)


def test_lbfgs_quadratic():
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((6, 6))
    A = Q @ Q.T + 0.1 * np.eye(6)
    b = rng.standard_normal(6)
    xs = np.linalg.solve(A, b)
    x, hist = lbfgs(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(6), iters=60)
    np.testing.assert_allclose(x, xs, rtol=1e-8, atol=1e-10)
    assert all(h1 <= h0 for h0, h1 in zip(hist, hist[1:]))


def test_lbfgs_steps_back_from_refused_points():
    # the function refuses x < -1 (as the engine refuses parameters
    # without a smooth cutoff): the search must back off and still converge
    def fun(x):
        if x[0] < -1:
            raise RuntimeError("refused")
        return float((x[0] + 0.9) ** 2 * 50 + x[1] ** 2), np.array([100 * (x[0] + 0.9), 2 * x[1]])
    x, _ = lbfgs(fun, np.array([3.0, 1.0]), iters=50)
    np.testing.assert_allclose(x, [-0.9, 0.0], atol=1e-8)


class ToyForceMatch:
    """pair forces of 4 LJ type pairs with Coulomb terms, at distances
    spread from 0.88 sigma outwards: F = 24 eps / r (2 (s/r)^12 - (s/r)^6)
    + k q_a q_b / r^2, weighted per type pair"""
    npl, npar, exp6 = 8, 11, False

    def __init__(self, seed=1, npair=6000):
        rng = np.random.default_rng(seed)
        self.t = rng.integers(0, 4, npair)
        self.r = np.minimum(SIG[self.t] * (0.88 + rng.exponential(0.25, npair)), 10.0)
        self.w = np.sqrt(np.array([1.0, 0.2, 0.2, 0.02])[self.t])
        self.target = self.forces(THT)

    def forces(self, th):
        s, e = th[0:8:2][self.t], th[1:8:2][self.t]
        qO, qNa, qCl = th[8:]
        qa = np.array([qO, qNa, qCl, qNa])[self.t]
        qb = np.array([qO, qO, qO, qCl])[self.t]
        sr6 = (s / self.r) ** 6
        return self.w * (24 * e / self.r * (2 * sr6 * sr6 - sr6) + 138.9 * qa * qb / self.r ** 2)

    def jac(self, th):
        J = np.empty((len(self.r), 11))
        for k in range(11):
            d = np.zeros(11)
            d[k] = 1e-6 * abs(th[k])
            J[:, k] = (self.forces(th + d) - self.forces(th - d)) / (2 * d[k])
        return J

    def loss_and_grad(self, th):
        d = self.forces(th) - self.target
        return float(d @ d), 2 * self.jac(th).T @ d

    def preconditioner(self, th, fv):
        J = self.jac(th) @ fv.jac(th)
        return np.linalg.cholesky(J.T @ J).T


def test_fit_recovers_toy_targets():
    fm = ToyForceMatch()
    th, hist = fit_lbfgs(fm, TH0, iters=40)
    assert hist[-1] < 1e-12 * hist[0]
    np.testing.assert_allclose(th, THT, rtol=1e-6)


def test_lbfgs_stops_at_rounding_floor():
    # a loss with a floor of rounding noise: the fit stops within a few
    # iterations of reaching it instead of spending every iteration there
    rng = np.random.default_rng(0)

    def fun(x):
        return float((x[0] - 1) ** 2 + 1 + 1e-12 * rng.random()), np.array([2 * (x[0] - 1)])
    x, hist = lbfgs(fun, np.array([5.0]), iters=100)
    assert abs(x[0] - 1) < 1e-4
    assert len(hist) < 15
