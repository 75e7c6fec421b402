"""Pins the leave-one-out oracle (oracle/objectives.py) before it checks the
GPU path: the reference's golden LOO value and gradient, brute-force
leave-one-out universal kriging, and finite differences."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err
from oracle import objectives as OB

KERNELS = ["gauss", "exp", "matern3_2", "matern5_2"]


def _data1():
    X = np.genfromtxt(os.path.join(GOLDEN, "data1-scal-X.csv"), delimiter=",").reshape(-1, 1)
    y = np.genfromtxt(os.path.join(GOLDEN, "data1-scal-y.csv"), delimiter=",")
    loo = float(np.genfromtxt(os.path.join(GOLDEN, "data1-scal-result-leaveOneOut.csv"), delimiter=","))
    g = np.atleast_1d(np.genfromtxt(os.path.join(GOLDEN, "data1-scal-result-leaveOneOutGrad.csv"), delimiter=","))
    return X, y, loo, g


def test_golden_loo_value_and_gradient():
    X, y, loo_ref, g_ref = _data1()
    loo, g, _, _ = OB.leave_one_out(X, y, "gauss", np.array([0.3]))
    assert rel_err(loo, loo_ref) < 1e-12
    assert rel_err(g, g_ref) < 1e-12


def test_golden_catches_a_dropped_trend_projection():
    # B without the - A'A term (simple instead of universal kriging) must fail
    X, y, loo_ref, _ = _data1()
    loo, _, _, _ = OB.leave_one_out(X, y, "gauss", np.array([0.3]), regmodel="none")
    assert rel_err(loo, loo_ref) > 1e-6


def _design(n, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(4 * X[:, 0]) + X[:, -1] ** 2 + 0.3 * X.sum(axis=1)
    return X, y


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["none", "constant", "linear"])
def test_loo_equals_bruteforce(kernel, regmodel):
    X, y = _design(25, 2, 3)
    theta = np.array([0.4, 0.6])
    loo, _, yhat, s2 = OB.leave_one_out(X, y, kernel, theta, regmodel=regmodel, want_grad=False)
    err, var = OB.leave_one_out_bruteforce(X, y, kernel, theta, regmodel=regmodel)
    np.testing.assert_allclose(y - yhat, err, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(s2, var, rtol=1e-8, atol=1e-12)
    assert rel_err(loo, np.mean(err ** 2)) < 1e-8


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["constant", "linear"])
def test_loo_gradient_vs_finite_differences(kernel, regmodel):
    X, y = _design(20, 3, 7)
    theta = np.array([0.5, 0.3, 0.8])
    _, g, _, _ = OB.leave_one_out(X, y, kernel, theta, regmodel=regmodel)
    fd = np.zeros(3)
    for k in range(3):
        h = 1e-6 * theta[k]
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fp = OB.leave_one_out(X, y, kernel, tp, regmodel=regmodel, want_grad=False)[0]
        fm = OB.leave_one_out(X, y, kernel, tm, regmodel=regmodel, want_grad=False)[0]
        fd[k] = (fp - fm) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9 * np.max(np.abs(fd)))


# --- log marginal posterior (Kriging::_logMargPost) -------------------------
# No reference fixture holds LMP values: the value is pinned by a closed form
# of the marginal likelihood (slogdet, no Cholesky) and a scaling identity,
# the gradient by finite differences of the value.

def _lml_closed_form(X, y, kernel, theta, regmodel):
    """-1/2 log|R| - 1/2 log|F'R^-1 F| - (n-p)/2 log(y'(R^-1 - R^-1 F (F'R^-1F)^-1 F'R^-1) y)"""
    from oracle import oracle as O
    n, d = X.shape
    R = O.corr_mat(kernel, X, X, theta)
    np.fill_diagonal(R, 1.0)
    F = O.trend_matrix(X, regmodel)
    p = F.shape[1]
    Ri = np.linalg.inv(R)
    M = F.T @ Ri @ F
    Q = Ri - Ri @ F @ np.linalg.solve(M, F.T @ Ri) if p else Ri
    ld_M = np.linalg.slogdet(M)[1] if p else 0.0
    lml = -0.5 * np.linalg.slogdet(R)[1] - 0.5 * ld_M - (n - p) / 2.0 * np.log(y @ Q @ y)
    a = 0.2
    b = (a + d) / n ** (1.0 / d)
    CL = (X.max(axis=0) - X.min(axis=0)) / n ** (1.0 / d)
    t = np.sum(CL / theta)
    return lml - b * t + a * np.log(t)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["none", "constant", "linear"])
def test_lmp_closed_form(kernel, regmodel):
    X, y = _design(30, 2, 4)
    theta = np.array([0.4, 0.7])
    lmp, _ = OB.log_marg_post(X, y, kernel, theta, regmodel=regmodel, want_grad=False)
    from oracle import oracle as O
    R = O.corr_mat(kernel, X, X, theta)
    np.fill_diagonal(R, 1.0)
    # the closed form inverts R explicitly: agreement to cond(R) eps
    tol = max(1e-12, 10 * np.linalg.cond(R) * np.finfo(float).eps)
    assert rel_err(lmp, _lml_closed_form(X, y, kernel, theta, regmodel)) < tol


def test_lmp_scaling_identity():
    # y -> c y moves S_2 by c^2: LMP drops by (n - p) log c
    X, y = _design(30, 2, 4)
    theta = np.array([0.4, 0.7])
    l1, _ = OB.log_marg_post(X, y, "matern5_2", theta, want_grad=False)
    l2, _ = OB.log_marg_post(X, 3.0 * y, "matern5_2", theta, want_grad=False)
    assert abs((l1 - l2) - (30 - 1) * np.log(3.0)) < 1e-9


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["constant", "linear"])
def test_lmp_gradient_vs_finite_differences(kernel, regmodel):
    X, y = _design(20, 3, 7)
    theta = np.array([0.5, 0.3, 0.8])
    _, g = OB.log_marg_post(X, y, kernel, theta, regmodel=regmodel)
    fd = np.zeros(3)
    for k in range(3):
        h = 1e-6 * theta[k]
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (OB.log_marg_post(X, y, kernel, tp, regmodel=regmodel, want_grad=False)[0]
                 - OB.log_marg_post(X, y, kernel, tm, regmodel=regmodel, want_grad=False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8 * np.max(np.abs(fd)))


@pytest.mark.parametrize("kernel", ["gauss", "matern3_2"])
def test_lmp_nugget_gradient_vs_finite_differences(kernel):
    X, y = _design(25, 2, 9)
    theta = np.array([0.4, 0.6])
    alpha = 0.9
    _, g = OB.log_marg_post(X, y, kernel, theta, alpha=alpha)
    fd = np.zeros(3)
    for k in range(3):
        tp, tm = theta.copy(), theta.copy()
        ap = am = alpha
        if k < 2:
            h = 1e-6 * theta[k]
            tp[k] += h
            tm[k] -= h
        else:
            h = 1e-7
            ap, am = alpha + h, alpha - h
        fd[k] = (OB.log_marg_post(X, y, kernel, tp, alpha=ap, want_grad=False)[0]
                 - OB.log_marg_post(X, y, kernel, tm, alpha=am, want_grad=False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8 * np.max(np.abs(fd)))
