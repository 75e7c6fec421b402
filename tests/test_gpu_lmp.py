"""Log marginal posterior on the GPU (Kriging.cpp:488-672, KrigingImpl.cpp:
887-921) against the CPU oracle (oracle/objectives.py, pinned by a closed
form and finite differences; no reference fixture holds LMP values).
Tolerances: value 1e-10 relative, gradient 1e-8 relative, scaled by cond(R)
for ill-conditioned cases."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import objectives as OB
from oracle import oracle as O

pytestmark = pytest.mark.gpu
KERNELS = ["gauss", "exp", "matern3_2", "matern5_2"]


def _design(n, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(4 * X[:, 0]) + X[:, -1] ** 2 + 0.3 * X.sum(axis=1) + 0.01 * rng.normal(size=n)
    return X, y


def _tols(kernel, X, theta, alpha=1.0):
    R = alpha * O.corr_mat(kernel, X, X, theta)
    np.fill_diagonal(R, 1.0)
    c = np.linalg.cond(R)
    return c, max(1e-10, 1e2 * c * np.finfo(float).eps), max(1e-8, 1e4 * c * np.finfo(float).eps)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["none", "constant", "linear"])
def test_lmp_vs_oracle(kernel, regmodel):
    import libkriging_amd as lk
    X, y = _design(150, 3, 13)
    theta = np.array([0.3, 0.5, 0.4])
    k = lk.Kriging(y, X, kernel, regmodel, False, "none", "LL", {"theta": np.atleast_2d(theta)})
    v, g = k.logMargPostFun(theta, True)
    vo, go = OB.log_marg_post(X, y, kernel, theta, regmodel=regmodel)
    c, tv, tg = _tols(kernel, X, theta)
    print(kernel, regmodel, "cond", c, rel_err(v, vo), rel_err(g, go))
    assert rel_err(v, vo) < tv
    assert rel_err(g, go) < tg


@pytest.mark.parametrize("kernel", ["gauss", "matern5_2"])
def test_lmp_nugget_vs_oracle(kernel):
    import libkriging_amd as lk
    X, y = _design(130, 2, 17)
    theta = np.array([0.3, 0.45])
    alpha = 0.95
    k = lk.Kriging(y, X, kernel, "constant", False, "none", "LL",
                   {"theta": np.atleast_2d(theta), "sigma2": 0.5, "nugget": 0.02}, "nugget")
    v, g = k.logMargPostFun(np.append(theta, alpha), True)
    vo, go = OB.log_marg_post(X, y, kernel, theta, alpha=alpha)
    c, tv, tg = _tols(kernel, X, theta, alpha)
    print(kernel, "cond", c, rel_err(v, vo), rel_err(g, go), g, go)
    assert g.shape == (3,)
    assert rel_err(v, vo) < tv
    assert rel_err(g, go) < tg


def test_lmp_fixed_sigma2_forward_differences():
    """sigma2 fixed: the reference's gradient is a forward difference of the
    objective (Kriging.cpp:634-646); both sides difference the same values"""
    import libkriging_amd as lk
    X, y = _design(100, 2, 3)
    theta = np.array([0.35, 0.5])
    k = lk.Kriging(y, X, "matern3_2", "constant", False, "none", "LL",
                   {"theta": np.atleast_2d(theta), "sigma2": 0.3, "is_sigma2_estim": False})
    v, g = k.logMargPostFun(theta, True)
    vo, go = OB.log_marg_post(X, y, "matern3_2", theta, est_sigma2=False, sigma2=0.3)
    assert rel_err(v, vo) < 1e-10
    np.testing.assert_allclose(g, go, rtol=1e-5, atol=1e-6 * np.max(np.abs(go)))


def test_lmp_large_design_gradient_fd():
    import libkriging_amd as lk
    X, y = _design(600, 4, 23)
    theta = np.array([0.3, 0.4, 0.5, 0.35])
    k = lk.Kriging(y, X, "matern5_2", "linear", False, "none", "LL", {"theta": np.atleast_2d(theta)})
    v, g = k.logMargPostFun(theta, True)
    vo, _ = OB.log_marg_post(X, y, "matern5_2", theta, regmodel="linear", want_grad=False)
    assert rel_err(v, vo) < 1e-10
    fd = np.zeros(4)
    for j in range(4):
        h = 1e-5 * theta[j]
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (k.logMargPostFun(tp)[0] - k.logMargPostFun(tm)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7 * np.max(np.abs(fd)))


def test_fit_lmp_objective():
    """fit(objective="LMP"): no worse than sampled points, batched == one
    evaluation per launch, sigma2 = SSE / (n - p) (Kriging.cpp:2192-2195)"""
    import libkriging_amd as lk
    X, y = _design(110, 2, 29)
    k = lk.Kriging(y, X, "matern5_2", "constant", False, "BFGS4", "LMP")
    assert k.objective() == "LMP"
    th = k.theta().reshape(-1)
    best = k.logMargPost()
    assert rel_err(best, k.logMargPostFun(th)[0]) < 1e-12
    rng = np.random.default_rng(4)
    for _ in range(20):
        t = np.exp(rng.uniform(np.log(0.05), np.log(2.0), size=2))
        assert best >= k.logMargPostFun(t)[0] - 1e-3 * abs(best)
    k2 = lk.Kriging("matern5_2")
    k2.set_batched_fit(False)
    k2.fit(y, X, "constant", False, "BFGS4", "LMP")
    np.testing.assert_allclose(k2.theta(), k.theta(), rtol=1e-9)
    kl = lk.Kriging(y, X, "matern5_2", "constant", False, "none", "LL", {"theta": np.atleast_2d(th)})
    n, p = 110, 1
    assert rel_err(k.sigma2(), kl.sigma2() * n / (n - p)) < 1e-10


def test_fit_lmp_nugget():
    import libkriging_amd as lk
    X, y = _design(90, 2, 31)
    k = lk.Kriging(y, X, "gauss", "constant", False, "BFGS2", "LMP", {}, "nugget")
    v, g = k.logMargPostFun(np.append(k.theta().reshape(-1), k.sigma2() / (k.sigma2() + k.nugget())), True)
    assert np.isfinite(v) and np.all(np.isfinite(g))
    assert k.sigma2() > 0 and k.nugget() >= 0
