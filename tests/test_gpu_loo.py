"""Leave-one-out objective on the GPU (Kriging.cpp:353-485, 1471-1486) against
the pinned CPU oracle (oracle/objectives.py) and the reference's golden
vectors.  Tolerances: value 1e-10 relative, gradient 1e-8 relative (scaled by
cond(R) for the ill-conditioned cases), as for the log-likelihood."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err
from oracle import objectives as OB
from oracle import oracle as O

pytestmark = pytest.mark.gpu
KERNELS = ["gauss", "exp", "matern3_2", "matern5_2"]


def _model(X, y, kernel, regmodel, theta):
    import libkriging_amd as lk
    return lk.Kriging(y, X, kernel, regmodel, False, "none", "LL", {"theta": np.atleast_2d(theta)})


def _design(n, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(4 * X[:, 0]) + X[:, -1] ** 2 + 0.3 * X.sum(axis=1) + 0.01 * rng.normal(size=n)
    return X, y


def test_golden_loo():
    X = np.genfromtxt(os.path.join(GOLDEN, "data1-scal-X.csv"), delimiter=",").reshape(-1, 1)
    y = np.genfromtxt(os.path.join(GOLDEN, "data1-scal-y.csv"), delimiter=",")
    loo_ref = float(np.genfromtxt(os.path.join(GOLDEN, "data1-scal-result-leaveOneOut.csv"), delimiter=","))
    g_ref = np.atleast_1d(np.genfromtxt(os.path.join(GOLDEN, "data1-scal-result-leaveOneOutGrad.csv"), delimiter=","))
    k = _model(X, y, "gauss", "constant", [0.3])
    loo, g = k.leaveOneOutFun(np.array([0.3]), True)
    print("golden", loo, loo_ref, g, g_ref)
    assert rel_err(loo, loo_ref) < 1e-10
    assert rel_err(g, g_ref) < 1e-8


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["none", "constant", "linear"])
def test_loo_vs_oracle(kernel, regmodel):
    # n = 150: three 64-row blocks, a ragged last one
    X, y = _design(150, 3, 11)
    theta = np.array([0.3, 0.5, 0.4])
    k = _model(X, y, kernel, regmodel, theta)
    loo, g = k.leaveOneOutFun(theta, True)
    loo_o, g_o, yhat_o, s2_o = OB.leave_one_out(X, y, kernel, theta, regmodel=regmodel)
    R = O.corr_mat(kernel, X, X, theta)
    np.fill_diagonal(R, 1.0)
    c = np.linalg.cond(R)
    tv = max(1e-10, 1e2 * c * np.finfo(float).eps)
    tg = max(1e-8, 1e4 * c * np.finfo(float).eps)
    print(kernel, regmodel, "cond", c, "rel loo", rel_err(loo, loo_o), "rel grad", rel_err(g, g_o))
    assert rel_err(loo, loo_o) < tv
    assert rel_err(g, g_o) < tg
    yhat, se = k.leaveOneOutVec(theta)
    np.testing.assert_allclose(yhat, yhat_o, rtol=tv, atol=tv * np.max(np.abs(y)))
    np.testing.assert_allclose(se, np.sqrt(s2_o * k.sigma2()), rtol=tv, atol=0)


def test_loo_larger_design():
    # n = 700 (11 blocks): value against the oracle, gradient against central
    # differences of the GPU value itself
    X, y = _design(700, 4, 5)
    theta = np.array([0.25, 0.35, 0.5, 0.3])
    k = _model(X, y, "matern5_2", "linear", theta)
    loo, g = k.leaveOneOutFun(theta, True)
    loo_o, _, _, _ = OB.leave_one_out(X, y, "matern5_2", theta, regmodel="linear", want_grad=False)
    assert rel_err(loo, loo_o) < 1e-9
    fd = np.zeros(4)
    for j in range(4):
        h = 1e-5 * theta[j]
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (k.leaveOneOutFun(tp)[0] - k.leaveOneOutFun(tm)[0]) / (2 * h)
    print("fd", g, fd)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7 * np.max(np.abs(fd)))


def test_fit_loo_objective():
    """fit(objective="LOO"): the optimum is no worse than sampled points, the
    batched and one-evaluation-per-launch fits agree, and the committed model
    reports its LOO"""
    import libkriging_amd as lk
    X, y = _design(120, 2, 21)
    k = lk.Kriging(y, X, "matern3_2", "constant", False, "BFGS4", "LOO")
    th = k.theta().reshape(-1)
    best = k.leaveOneOut()
    assert k.objective() == "LOO"
    assert rel_err(best, k.leaveOneOutFun(th)[0]) < 1e-12
    rng = np.random.default_rng(2)
    for _ in range(20):
        t = np.exp(rng.uniform(np.log(0.05), np.log(2.0), size=2))
        assert best <= k.leaveOneOutFun(t)[0] * (1 + 1e-3)
    k2 = lk.Kriging("matern3_2")
    k2.set_batched_fit(False)
    k2.fit(y, X, "constant", False, "BFGS4", "LOO")
    np.testing.assert_allclose(k2.theta(), k.theta(), rtol=1e-9)
    # sigma2 committed as SSE / n (Kriging.cpp:2190-2195), as for LL
    kl = _model(X, y, "matern3_2", "constant", th)
    assert rel_err(k.sigma2(), kl.sigma2()) < 1e-10


def test_loo_rejects_nugget_model():
    import libkriging_amd as lk
    X, y = _design(40, 2, 1)
    with pytest.raises((ValueError, RuntimeError), match="LOO objective not supported"):
        lk.Kriging(y, X, "gauss", "constant", False, "BFGS", "LOO", {}, "nugget")
