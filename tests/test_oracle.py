"""Pins the CPU oracle (oracle/lk_oracle.c) before it is trusted as the
checker of the GPU path:
  - the reference's golden vectors (tests/golden, from
    tests/references/data1-scal-*, data2-grad-*) at 1e-12 (the tolerance of
    bindings/Python/pylibkriging/tests/binding_consistency_test.py:11);
  - LAPACK through scipy (dpotrf, dtrcon) for the factorisation and the
    condition estimate arma::rcond uses;
  - closed forms and special cases of the four kernels;
  - central finite differences of the LL for every gradient component,
    including the nugget direction (the check of
    tests/NuggetKrigingLogLikTest.cpp:11-191);
  - an independent numpy formulation of the concentrated LL;
  - the jitter ladder (LinearAlgebra.cpp:68-98) on a singular matrix.
"""
import os

import numpy as np
import pytest
from scipy.linalg import lapack

from conftest import GOLDEN, rel_err
from oracle import oracle as O

KERNELS = ["gauss", "exp", "matern3_2", "matern5_2"]


def _load(prefix):
    X = np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-X.csv"), delimiter=",")
    y = np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-y.csv"), delimiter=",")
    llr = float(np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-result-logLikelihood.csv"), delimiter=","))
    gr = np.atleast_1d(np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-result-logLikelihoodGrad.csv"), delimiter=","))
    return X.reshape(len(y), -1), y, llr, gr


@pytest.mark.parametrize("prefix", ["data1-scal"] + [f"data2-grad-{i}" for i in range(1, 11)])
def test_golden_vectors(prefix):
    X, y, llr, gr = _load(prefix)
    P = O.OracleProblem(X, y, "gauss")
    ll, g, inc = P.log_likelihood(0.3 * np.ones(X.shape[1]))
    assert rel_err(ll, llr) < 1e-12
    assert rel_err(g, gr) < 1e-12
    assert inc == 0


def test_planted_mistake_is_caught_by_golden_vectors():
    # a dropped factor 2 in the log-det (a plausible slip) must fail the pin
    X, y, llr, _ = _load("data2-grad-5")
    P = O.OracleProblem(X, y, "gauss")
    ll, _, _ = P.log_likelihood(0.3 * np.ones(X.shape[1]))
    R = P.build_R(0.3 * np.ones(X.shape[1]))
    half_logdet = 0.5 * np.linalg.slogdet(R)[1]
    assert rel_err(ll + 0.5 * half_logdet * 2, llr) > 1e-6


def _numpy_ll(X, y, kernel, theta, F):
    """Independent formulation: dense numpy Cholesky / solves / GLS."""
    n = len(y)
    D = X[:, None, :] - X[None, :, :]
    u = np.abs(D) / theta
    if kernel == "gauss":
        R = np.exp(-0.5 * np.sum(u ** 2, axis=2))
    elif kernel == "exp":
        R = np.exp(-np.sum(u, axis=2))
    elif kernel == "matern3_2":
        v = np.sqrt(3) * u
        R = np.prod((1 + v) * np.exp(-v), axis=2)
    else:
        v = np.sqrt(5) * u
        R = np.prod((1 + v + v ** 2 / 3) * np.exp(-v), axis=2)
    np.fill_diagonal(R, 1.0)
    L = np.linalg.cholesky(R)
    Fs = np.linalg.solve(L, F)
    ys = np.linalg.solve(L, y)
    beta = np.linalg.lstsq(Fs, ys, rcond=None)[0]
    E = np.linalg.solve(L, y - F @ beta)
    s2 = E @ E / n
    return -0.5 * (n * np.log(2 * np.pi * s2) + 2 * np.sum(np.log(np.diag(L))) + n)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("regmodel", ["constant", "linear"])
def test_ll_matches_numpy_formulation(kernel, regmodel):
    rng = np.random.default_rng(3)
    X = rng.uniform(size=(40, 3))
    y = np.sin(3 * X).sum(axis=1)
    theta = np.array([0.3, 0.5, 0.4])
    P = O.OracleProblem(X, y, kernel, regmodel)
    ll, _, _ = P.log_likelihood(theta, want_grad=False)
    ll_np = _numpy_ll(X, y, kernel, theta, O.trend_matrix(X, regmodel))
    assert rel_err(ll, ll_np) < 1e-11


@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_special_values(kernel):
    th = np.array([0.3, 0.7])
    assert O.cov(kernel, np.zeros(2), th) == 1.0
    assert np.all(O.dlncov_dtheta(kernel, np.zeros(2), th) == 0.0)
    # separable kernels factor over dimensions
    dx = np.array([0.11, -0.23])
    c = O.cov(kernel, dx, th)
    c1 = O.cov(kernel, dx[:1], th[:1]) * O.cov(kernel, dx[1:], th[1:])
    assert abs(c - c1) < 1e-15
    # symmetric in dx
    assert O.cov(kernel, -dx, th) == c


def test_kernel_closed_forms():
    th = np.array([0.4])
    for h in [0.05, 0.3, 1.2]:
        dx = np.array([h])
        u = h / 0.4
        assert abs(O.cov("gauss", dx, th) - np.exp(-0.5 * u * u)) < 1e-15
        assert abs(O.cov("exp", dx, th) - np.exp(-u)) < 1e-15
        v3 = np.sqrt(3) * u
        assert abs(O.cov("matern3_2", dx, th) - (1 + v3) * np.exp(-v3)) < 1e-14
        v5 = np.sqrt(5) * u
        assert abs(O.cov("matern5_2", dx, th) - (1 + v5 + v5 * v5 / 3) * np.exp(-v5)) < 1e-14


@pytest.mark.parametrize("kernel", KERNELS)
def test_dlncov_by_finite_differences(kernel):
    dx = np.array([0.13, -0.4, 0.05])
    th = np.array([0.3, 0.8, 0.2])
    g = O.dlncov_dtheta(kernel, dx, th)
    for k in range(3):
        h = 1e-6 * th[k]
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (np.log(O.cov(kernel, dx, tp)) - np.log(O.cov(kernel, dx, tm))) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * max(1.0, abs(fd))


def _fd_grad(P, gamma, rel=1e-6):
    g = np.zeros_like(gamma)
    for k in range(len(gamma)):
        h = rel * abs(gamma[k])
        gp, gm = gamma.copy(), gamma.copy()
        gp[k] += h
        gm[k] -= h
        g[k] = (P.log_likelihood(gp, False)[0] - P.log_likelihood(gm, False)[0]) / (2 * h)
    return g


@pytest.mark.parametrize("kernel", KERNELS)
def test_gradient_by_finite_differences(kernel):
    rng = np.random.default_rng(8)
    X = rng.uniform(size=(30, 2))
    y = np.cos(4 * X[:, 0]) + X[:, 1]
    P = O.OracleProblem(X, y, kernel)
    gamma = np.array([0.3, 0.45])
    _, g, _ = P.log_likelihood(gamma)
    assert rel_err(g, _fd_grad(P, gamma)) < 1e-5


@pytest.mark.parametrize("est_nugget", [True, False])
def test_nugget_gradient_by_finite_differences(est_nugget):
    rng = np.random.default_rng(9)
    X = rng.uniform(size=(30, 2))
    y = np.cos(4 * X[:, 0]) + X[:, 1] + rng.normal(scale=0.05, size=30)
    P = O.OracleProblem(X, y, "matern5_2", "constant", "nugget", est_nugget=est_nugget, nugget=0.01)
    gamma = np.array([0.3, 0.45, 0.9])
    _, g, _ = P.log_likelihood(gamma)
    fd = _fd_grad(P, gamma, rel=1e-7)
    assert rel_err(g, fd) < 1e-5


def test_fixed_sigma2_gradient_by_finite_differences():
    rng = np.random.default_rng(10)
    X = rng.uniform(size=(25, 2))
    y = np.sin(5 * X[:, 0])
    P = O.OracleProblem(X, y, "gauss", est_sigma2=False, sigma2=0.7)
    gamma = np.array([0.2, 0.35])
    _, g, _ = P.log_likelihood(gamma)
    assert rel_err(g, _fd_grad(P, gamma)) < 1e-5


def test_potrf_matches_lapack():
    rng = np.random.default_rng(11)
    A = rng.normal(size=(50, 50))
    A = A @ A.T + 50 * np.eye(50)
    L, info = O.potrf_lower(A)
    Lr, inf2 = lapack.dpotrf(A, lower=1)
    assert info == 0 and inf2 == 0
    assert np.abs(L - np.tril(Lr)).max() < 1e-12
    # a non-SPD matrix reports the failing pivot like dpotrf
    B = A.copy()
    B[10, 10] = -1.0
    _, info = O.potrf_lower(B)
    _, info2 = lapack.dpotrf(B, lower=1)
    assert info == info2 > 0


def test_rcond_estimate_matches_lapack_dtrcon():
    # arma::rcond on a triangular factor calls LAPACK dtrcon (1-norm)
    rng = np.random.default_rng(12)
    for n in [1, 2, 7, 40, 120]:
        X = rng.uniform(size=(n, 2))
        R = np.exp(-0.5 * np.sum((X[:, None, :] - X[None, :, :]) ** 2, axis=2) / 0.3 ** 2) + 1e-8 * np.eye(n)
        L = np.linalg.cholesky(R)
        rc = O.trcon_lower(L)
        rc_ref, info = lapack.dtrcon(L, norm="1", uplo="L", diag="N")
        assert info == 0
        assert abs(rc - rc_ref) <= 1e-12 * rc_ref


def test_jitter_ladder_on_singular_matrix():
    # two identical design points: R singular; the ladder adds
    # num_nugget * sum_{k < inc} 10^k to the diagonal
    X = np.array([[0.1, 0.2], [0.4, 0.9], [0.1, 0.2], [0.7, 0.3]])
    P = O.OracleProblem(X, np.arange(4.0), "gauss")
    R = P.build_R(np.array([0.3, 0.3]))
    L, inc, rc = O.safe_chol_lower(R)
    assert inc >= 1
    jit = sum(1e-10 * 10.0 ** q for q in range(inc))
    Lr = np.linalg.cholesky(R + jit * np.eye(4))
    assert np.abs(L - Lr).max() < 1e-10
    assert rc >= 1e-18
    # with a non-positive unit the ladder must fail
    with pytest.raises(RuntimeError):
        O.safe_chol_lower(R, O.default_cfg(num_nugget=0.0))


def test_theta_bounds_numpy_restatement():
    rng = np.random.default_rng(14)
    X = rng.uniform(size=(25, 3))
    y = np.sin(4 * X).sum(axis=1)
    lo, up = O.theta_bounds(X, y)
    maxdx = X.max(0) - X.min(0)
    D = X[:, None, :] - X[None, :, :]
    dy2 = (y[:, None] - y[None, :]) ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        w = dy2 / np.sum(D ** 2, axis=2)
    w[np.isnan(w)] = 0.0
    w /= w.sum()
    steep = np.einsum("ij,ijk->k", w, np.abs(D))
    lo_r = np.minimum(np.maximum(0.02 * maxdx, 0.02 * steep), 10 * maxdx)
    assert np.allclose(lo, lo_r, rtol=1e-12)
    assert np.allclose(up, 10 * maxdx, rtol=1e-12)


def test_predict_interpolates_and_matches_numpy():
    rng = np.random.default_rng(15)
    X = rng.uniform(size=(30, 2))
    y = np.sin(5 * X[:, 0]) + X[:, 1]
    P = O.OracleProblem(X, y, "matern5_2")
    th = np.array([0.3, 0.5])
    m, s = P.predict(th, 1.0, 1.3, X)
    assert np.allclose(m, y, atol=1e-9)
    assert np.all(s < 1e-6)
    Xn = rng.uniform(size=(10, 2))
    m, s = P.predict(th, 1.0, 1.3, Xn)
    # simple-kriging-with-GLS-trend formulas in numpy
    def k(A, B):
        v = np.sqrt(5) * np.abs(A[:, None, :] - B[None, :, :]) / th
        return np.prod((1 + v + v ** 2 / 3) * np.exp(-v), axis=2)
    R = k(X, X)
    F = np.ones((30, 1))
    Ri = np.linalg.inv(R)
    beta = np.linalg.solve(F.T @ Ri @ F, F.T @ Ri @ y)
    r = k(X, Xn)
    mean = beta[0] + r.T @ Ri @ (y - F @ beta)
    u = 1 - F.T @ Ri @ r
    var = 1 - np.sum(r * (Ri @ r), axis=0) + (u ** 2).ravel() / (F.T @ Ri @ F).item()
    assert np.allclose(m, mean, rtol=1e-9)
    assert np.allclose(s, np.sqrt(1.3 * np.maximum(var, 0)), rtol=1e-6)


def _numpy_corr(X, kernel, theta):
    D = np.abs(X[:, None, :] - X[None, :, :]) / theta
    if kernel == "gauss":
        C = np.exp(-0.5 * np.sum(D ** 2, axis=2))
    elif kernel == "exp":
        C = np.exp(-np.sum(D, axis=2))
    elif kernel == "matern3_2":
        v = np.sqrt(3) * D
        C = np.prod((1 + v) * np.exp(-v), axis=2)
    else:
        v = np.sqrt(5) * D
        C = np.prod((1 + v + v ** 2 / 3) * np.exp(-v), axis=2)
    np.fill_diagonal(C, 1.0)
    return C


def _gauss_loglik(y, F, K):
    """log N(y; F beta, K) with beta the GLS estimate under K: a dense
    scipy formulation that knows nothing of alpha or of the model's branches"""
    from scipy.linalg import cho_factor, cho_solve
    cf = cho_factor(K, lower=True)
    KiF = cho_solve(cf, F)
    beta = np.linalg.solve(F.T @ KiF, KiF.T @ y)
    r = y - F @ beta
    n = len(y)
    return -0.5 * (n * np.log(2 * np.pi) + 2 * np.sum(np.log(np.diag(cf[0]))) + r @ cho_solve(cf, r))


# (est_sigma2, est_nugget): the nugget value branches of Kriging.cpp:247-268.
# For each, the covariance the model implies, from the parameter values alone:
#   sigma2 fixed, nugget estimated: tau2 = sigma2 (1 - alpha) / alpha
#   sigma2 estimated, nugget fixed: sigma2 = tau2 alpha / (1 - alpha)
#   both fixed: alpha = sigma2 / (sigma2 + tau2) (Kriging.cpp:241-242)
# and the LL is the Gaussian log-density of y under K = sigma2 C + tau2 I
@pytest.mark.parametrize("kernel", ["gauss", "matern5_2"])
@pytest.mark.parametrize("branch", ["sigma2_fixed", "nugget_fixed", "both_fixed"])
def test_nugget_value_branches_match_gaussian_density(kernel, branch):
    rng = np.random.default_rng(21)
    X = rng.uniform(size=(35, 2))
    y = np.cos(4 * X[:, 0]) + X[:, 1] + rng.normal(scale=0.1, size=35)
    theta, alpha, s2, tau2 = np.array([0.3, 0.45]), 0.8, 1.7, 0.05
    C = _numpy_corr(X, kernel, theta)
    F = np.ones((35, 1))
    if branch == "sigma2_fixed":
        P = O.OracleProblem(X, y, kernel, "constant", "nugget", est_sigma2=False, sigma2=s2)
        ll, _, _ = P.log_likelihood(np.array([*theta, alpha]), want_grad=False)
        K = s2 * C + s2 * (1 - alpha) / alpha * np.eye(35)
    elif branch == "nugget_fixed":
        P = O.OracleProblem(X, y, kernel, "constant", "nugget", est_nugget=False, nugget=tau2)
        ll, _, _ = P.log_likelihood(np.array([*theta, alpha]), want_grad=False)
        K = tau2 * alpha / (1 - alpha) * C + tau2 * np.eye(35)
    else:
        P = O.OracleProblem(X, y, kernel, "constant", "nugget", est_sigma2=False, sigma2=s2, est_nugget=False,
                            nugget=tau2)
        ll, _, _ = P.log_likelihood(theta, want_grad=False)
        K = s2 * C + tau2 * np.eye(35)
    assert rel_err(ll, _gauss_loglik(y, F, K)) < 1e-11


def test_nugget_branch_pin_catches_swapped_variance():
    # a plausible slip -- sigma2 = tau2 (1 - alpha) / alpha instead of
    # tau2 alpha / (1 - alpha) -- changes the density the pin computes
    rng = np.random.default_rng(21)
    X = rng.uniform(size=(35, 2))
    y = np.cos(4 * X[:, 0]) + X[:, 1] + rng.normal(scale=0.1, size=35)
    theta, alpha, tau2 = np.array([0.3, 0.45]), 0.8, 0.05
    C = _numpy_corr(X, "gauss", theta)
    P = O.OracleProblem(X, y, "gauss", "constant", "nugget", est_nugget=False, nugget=tau2)
    ll, _, _ = P.log_likelihood(np.array([*theta, alpha]), want_grad=False)
    wrong = _gauss_loglik(y, np.ones((35, 1)), tau2 * (1 - alpha) / alpha * C + tau2 * np.eye(35))
    assert rel_err(ll, wrong) > 1e-3


def test_noise_free_fixed_sigma2_matches_gaussian_density():
    rng = np.random.default_rng(22)
    X = rng.uniform(size=(30, 3))
    y = np.sin(3 * X).sum(axis=1)
    theta = np.array([0.4, 0.3, 0.6])
    P = O.OracleProblem(X, y, "exp", "linear", est_sigma2=False, sigma2=0.8)
    ll, _, _ = P.log_likelihood(theta, want_grad=False)
    assert rel_err(ll, _gauss_loglik(y, O.trend_matrix(X, "linear"), 0.8 * _numpy_corr(X, "exp", theta))) < 1e-11


def test_lapack_route_matches_restated_loops():
    """the oracle's optional LAPACK route (dpotrf / dtrtrs / dtrcon from
    scipy, the CPU baseline's arithmetic) equals the restated loops on the
    reference's golden vectors and on a nugget / linear-trend case"""
    try:
        for prefix in ["data1-scal", "data2-grad-7"]:
            X, y, llr, gr = _load(prefix)
            P = O.OracleProblem(X, y, "gauss")
            O.use_lapack(True)
            ll, g, inc = P.log_likelihood(0.3 * np.ones(X.shape[1]))
            O.use_lapack(False)
            assert rel_err(ll, llr) < 1e-12 and rel_err(g, gr) < 1e-12 and inc == 0
        rng = np.random.default_rng(31)
        X = rng.uniform(size=(120, 3))
        y = np.sin(3 * X).sum(axis=1)
        P = O.OracleProblem(X, y, "matern5_2", "linear", "nugget")
        gam = np.array([0.3, 0.4, 0.5, 0.9])
        a = P.log_likelihood(gam)
        O.use_lapack(True)
        b = P.log_likelihood(gam)
        L, info = O.potrf_lower(P.build_R(gam[:3]) + np.eye(120))
        rc = O.trcon_lower(np.tril(L))
        O.use_lapack(False)
        assert rel_err(a[0], b[0]) < 1e-13 and rel_err(a[1], b[1]) < 1e-12 and a[2] == b[2]
        assert info == 0 and rel_err(rc, lapack.dtrcon(np.tril(L), norm="1", uplo="L")[0]) < 1e-12
    finally:
        O.use_lapack(False)


def _spd(n, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, -np.log10(cond), n)
    return (Q * ev) @ Q.T


@pytest.mark.parametrize("n,no", [(7, 1), (40, 13), (90, 64), (130, 129)])
def test_chol_block_equals_the_unique_factor(n, no):
    # the Cholesky factor with a positive diagonal is unique: given the true
    # factor of the leading block, the block update must return chol(C)
    C = _spd(n, 10 + n)
    Loo = np.linalg.cholesky(C[:no, :no])
    L, path = O.chol_block(C, Loo)
    assert path == "block"
    assert np.allclose(L, np.tril(L))
    assert np.max(np.abs(L - np.linalg.cholesky(C))) < 1e-12 * np.max(np.abs(L))
    assert np.max(np.abs(L @ L.T - C)) < 1e-12


def test_chol_block_uses_the_given_factor():
    # a Loo that is not the factor of C_oo enters through L_ou = C_uo Loo^-T
    # and the Schur complement (scipy's triangular solve and numpy's
    # Cholesky as the independent route); C_oo itself is never read
    from scipy.linalg import solve_triangular
    n, no = 50, 20
    C = _spd(n, 3)
    Loo = 2.0 * np.linalg.cholesky(C[:no, :no])
    L, path = O.chol_block(C, Loo)
    assert path == "block"
    Lou = solve_triangular(Loo, C[:no, no:], lower=True)
    ref = np.zeros((n, n))
    ref[:no, :no] = Loo
    ref[no:, :no] = Lou.T
    ref[no:, no:] = np.linalg.cholesky(C[no:, no:] - Lou.T @ Lou)
    assert np.max(np.abs(L - ref)) < 1e-12
    Cb = C.copy()
    Cb[:no, :no] = 0.0
    assert np.array_equal(O.chol_block(Cb, Loo)[0], L)


def test_chol_block_falls_back_to_the_full_factor():
    # Loo = chol(C_oo) / 10 makes C_uu - L_ou L_ou' strongly indefinite:
    # the ladder on it is exhausted (its jitter totals ~1.1) and the result
    # is the factor of C from scratch (LinearAlgebra.cpp:286-296)
    n, no = 30, 12
    C = _spd(n, 5, cond=10.0)
    Loo = 0.1 * np.linalg.cholesky(C[:no, :no])
    L, path = O.chol_block(C, Loo)
    assert path == "full"
    assert np.max(np.abs(L - np.linalg.cholesky(C))) < 1e-12
