"""predict's covariance and derivative outputs (KrigingImpl.cpp:245-300).

cov is checked against the universal-kriging covariance written with dense
numpy solves (no L^-1, no R*):
    Sigma = s (R_nn - r' R^-1 r + E (F' R^-1 F)^-1 E'),  E = F_n - r' R^-1 F,
and its diagonal against the stdev output; the derivatives against central
finite differences of the GPU mean and stdev themselves."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _model(kernel, regmodel, noise, n=150, d=3, seed=5):
    import libkriging_amd as lk
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(3 * X[:, 0]) + X[:, 1] ** 2 - 0.5 * X[:, 2] + 0.05 * rng.normal(size=n)
    params = {"theta": np.array([[0.35, 0.45, 0.3]])}
    if noise == "nugget":
        params.update({"sigma2": 0.8, "nugget": 0.01, "is_sigma2_estim": False, "is_nugget_estim": False})
    k = lk.Kriging(y, X, kernel, regmodel, False, "none", "LL", params, noise)
    return k, X, y


def _trend(X, regmodel):
    return O.trend_matrix(X, regmodel)


def _fd(k, Xn, h=1e-4):
    """fourth-order central differences of the GPU mean and variance (stdev^2,
    smooth also where stdev -> 0 at design points): truncation O(h^4), and a
    step large enough that the predictor's own rounding (which grows with
    cond(R)) stays far below the tolerance"""
    m, d = Xn.shape
    fd_m, fd_s = np.zeros((m, d)), np.zeros((m, d))
    for j in range(d):
        vals = []
        for t in (2, 1, -1, -2):
            Xs = Xn.copy()
            Xs[:, j] += t * h
            mu, sd, _, _, _ = k.predict(Xs, True)
            vals.append((mu, sd))
        w = (-1.0, 8.0, -8.0, 1.0)
        fd_m[:, j] = sum(wi * v[0] for wi, v in zip(w, vals)) / (12 * h)
        fd_s[:, j] = sum(wi * v[1] ** 2 for wi, v in zip(w, vals)) / (12 * h)
    return fd_m, fd_s


@pytest.mark.parametrize("kernel,regmodel,noise", [("matern5_2", "constant", None), ("gauss", "linear", None),
                                                   ("matern3_2", "constant", "nugget"), ("exp", "linear", "nugget")])
def test_predict_cov_vs_dense(kernel, regmodel, noise):
    k, X, y = _model(kernel, regmodel, noise)
    rng = np.random.default_rng(8)
    Xn = rng.uniform(size=(70, 3))
    mean, sd, cov, _, _ = k.predict(Xn, True, True, False)
    th = k.theta()
    if noise == "nugget":
        alpha = k.sigma2() / (k.sigma2() + k.nugget())
        vs = k.sigma2() / alpha
    else:
        alpha, vs = 1.0, k.sigma2()
    R = alpha * O.corr_mat(kernel, X, X, th)
    np.fill_diagonal(R, 1.0)
    r = alpha * O.corr_mat(kernel, X, Xn, th)
    Rnn = alpha * O.corr_mat(kernel, Xn, Xn, th)
    np.fill_diagonal(Rnn, 1.0)
    F, Fn = _trend(X, regmodel), _trend(Xn, regmodel)
    Ri_r = np.linalg.solve(R, r)
    Ri_F = np.linalg.solve(R, F)
    E = Fn - r.T @ Ri_F
    S = vs * (Rnn - r.T @ Ri_r + E @ np.linalg.solve(F.T @ Ri_F, E.T))
    c = np.linalg.cond(R)
    tol = max(1e-9, 1e3 * c * np.finfo(float).eps)
    print(kernel, regmodel, noise, "cond", c, "max |dS|", np.max(np.abs(cov - S)), "scale", vs)
    assert cov.shape == (70, 70)
    assert np.max(np.abs(cov - S)) <= tol * vs
    assert np.max(np.abs(cov - cov.T)) <= tol * vs
    # diag(cov) and stdev^2 are both 1 - |r*|^2 + |E*|^2 but summed in other
    # orders: they agree to the cancellation error of that difference
    pos = np.diag(S) > 1e-8 * vs
    np.testing.assert_allclose(np.diag(cov)[pos], sd[pos] ** 2, rtol=1e-9, atol=tol * vs)


@pytest.mark.parametrize("kernel,regmodel,noise", [("matern5_2", "constant", None), ("gauss", "linear", None),
                                                   ("matern3_2", "linear", "nugget"), ("exp", "constant", None)])
def test_predict_derivatives_fd(kernel, regmodel, noise):
    k, X, y = _model(kernel, regmodel, noise)
    rng = np.random.default_rng(9)
    Xn = rng.uniform(0.05, 0.95, size=(25, 3))
    mean, sd, _, dm, ds = k.predict(Xn, True, False, True)
    assert dm.shape == (25, 3) and ds.shape == (25, 3)
    fd_m, fd_s = _fd(k, Xn)
    ds = 2 * sd[:, None] * ds  # d stdev^2 / dx
    # the exp kernel has a kink in x_k wherever x_k equals a design point's
    # coordinate: keep the (point, coordinate) pairs whose stencil holds none
    gap = np.min(np.abs(Xn[:, None, :] - X[None, :, :]), axis=1)
    ok = gap > 3e-4 if kernel == "exp" else np.ones_like(gap, dtype=bool)
    assert ok.sum() > 0.8 * ok.size
    scale_m = np.max(np.abs(fd_m))
    scale_s = np.max(np.abs(fd_s))
    em = np.max(np.abs(dm - fd_m)[ok])
    es = np.max(np.abs(ds - fd_s)[ok])
    print(kernel, regmodel, noise, em / scale_m, es / scale_s)
    assert em <= 1e-5 * scale_m
    assert es <= 1e-5 * scale_s


def test_predict_deriv_normalised_two_chunks():
    """normalize=True (the chain rule of KrigingImpl.cpp:289-293) and more
    points than one prediction chunk (8192): against finite differences"""
    import libkriging_amd as lk
    rng = np.random.default_rng(12)
    X = rng.uniform(size=(120, 2)) * np.array([3.0, 0.5]) + np.array([1.0, -2.0])
    y = 10.0 + np.sin(X[:, 0]) * np.exp(X[:, 1])
    # theta in normalised units; cond(R) ~ 2e6 keeps the differences clean
    k = lk.Kriging(y, X, "gauss", "constant", True, "none", "LL", {"theta": np.array([[0.08, 0.1]])})
    Xn = rng.uniform(size=(8500, 2)) * np.array([3.0, 0.5]) + np.array([1.0, -2.0])
    mean, sd, _, dm, ds = k.predict(Xn, True, False, True)
    fd_m, fd_s = _fd(k, Xn)
    ds = 2 * sd[:, None] * ds
    for j in range(2):
        assert np.max(np.abs(dm[:, j] - fd_m[:, j])) <= 1e-5 * np.max(np.abs(dm[:, j]))
        assert np.max(np.abs(ds[:, j] - fd_s[:, j])) <= 1e-5 * np.max(np.abs(ds[:, j]))


def test_predict_cov_limit():
    """m x m covariance over at most 32768 points per call (8 GiB); the
    output array is allocated lazily, so asking for one more is cheap"""
    k, X, y = _model("gauss", "constant", None, n=40)
    with pytest.raises(RuntimeError, match="32768"):
        k.predict(np.random.default_rng(1).uniform(size=(32769, 3)), True, True, False)


def test_predict_cov_two_tiles_rows_and_stdev_only_agree():
    """m not a multiple of the 64-point tile, several tiles per side: cov's
    diagonal against the stdev-only call, the mean unchanged by the flags"""
    k, X, y = _model("matern5_2", "constant", None)
    Xn = np.random.default_rng(3).uniform(size=(200, 3))
    mean, sd, cov, _, _ = k.predict(Xn, True, True, False)
    mean2, sd2, _, _, _ = k.predict(Xn, True)
    np.testing.assert_array_equal(mean, mean2)
    np.testing.assert_array_equal(sd, sd2)
    np.testing.assert_allclose(np.diag(cov), sd ** 2, rtol=1e-9, atol=1e-12 * k.sigma2())
    np.testing.assert_allclose(cov, cov.T, rtol=0, atol=1e-13 * k.sigma2())
