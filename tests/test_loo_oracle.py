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
