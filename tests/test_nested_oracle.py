"""Pins of the NestedKriging oracle (oracle/nested.py) by properties the
mathematics fixes, independent of the restatement's own formulas:
  - singleton groups: the NK predictor is the BLUP on (M_g) = (affine
    functions of single y_g), i.e. full simple kriging (dense solve);
  - one group: NK = that group's simple-kriging predictor;
  - NK interpolates the design (NestedKrigingTest.cpp:63-75);
  - PoE / gPoE / BCM with one expert return the expert; rBCM keeps its mean
    (NestedKrigingTest.cpp:44-58);
  - var <= sigma2, far field -> (beta0, sigma2) (NestedKrigingTest.cpp:139-154).
A planted mistake (dropping the cross terms, a transposed W) fails the first pin.
"""
import numpy as np
import pytest

from oracle import nested as NKO
from oracle import oracle as O


def f2d(X):
    return np.sin(3.0 * X[:, 0]) + np.cos(5.0 * X[:, 1]) + X[:, 0] * X[:, 1]


@pytest.fixture
def design():
    rng = np.random.default_rng(11)
    X = rng.uniform(size=(30, 2))
    return X, f2d(X), rng.uniform(size=(12, 2))


@pytest.mark.parametrize("kernel", ["gauss", "matern5_2", "exp"])
def test_singleton_groups_equal_full_simple_kriging(design, kernel):
    """without jitter, M_g = beta0 + k_g (y_g - beta0) is an invertible diagonal
    map of y, so the BLUP on M is the BLUP on y: exact full simple kriging"""
    X, y, Xn = design
    theta = np.array([0.25, 0.3])
    groups = [np.array([i]) for i in range(X.shape[0])]
    m_nk, s_nk = NKO.predict_nk(X, y, groups, theta, 1.3, 0.4, kernel, Xn, jitter=0.0)
    m_sk, s_sk = NKO.simple_kriging(X, y, theta, 1.3, 0.4, kernel, Xn)
    c = np.linalg.cond(O.corr_mat(kernel, X, X, theta))
    tol = max(1e-10, 1e3 * c * np.finfo(float).eps)
    assert np.max(np.abs(m_nk - m_sk)) <= tol * np.max(np.abs(y))
    assert np.max(np.abs(s_nk ** 2 - s_sk ** 2)) <= tol * 1.3


def test_planted_missing_cross_terms_fails(design):
    """the singleton-group pin detects an aggregation without the off-diagonal
    covariances (the PoE-like mistake)"""
    X, y, Xn = design
    theta = np.array([0.25, 0.3])
    groups = [np.array([i]) for i in range(X.shape[0])]
    Ls, alphas = NKO.precompute_nk(X, y, groups, theta, 0.4, "gauss", jitter=0.0)
    p = len(groups)
    M = np.array([0.4 + O.corr_mat("gauss", X[g], Xn, theta).T @ alphas[k] for k, g in enumerate(groups)])
    Kd = np.array([np.sum((np.linalg.solve(Ls[k], O.corr_mat("gauss", X[g], Xn, theta))) ** 2, axis=0)
                   for k, g in enumerate(groups)])
    wrong = np.array([0.4 + np.linalg.solve(np.diag(Kd[:, t]), Kd[:, t]) @ (M[:, t] - 0.4)
                      for t in range(Xn.shape[0])])
    m_sk, _ = NKO.simple_kriging(X, y, theta, 1.3, 0.4, "gauss", Xn)
    assert p == X.shape[0]
    assert np.max(np.abs(wrong - m_sk)) > 1e-3


def test_one_group_is_simple_kriging(design):
    X, y, Xn = design
    theta = np.array([0.3, 0.25])
    m_nk, s_nk = NKO.predict_nk(X, y, [np.arange(X.shape[0])], theta, 2.0, -0.1, "matern5_2", Xn)
    m_sk, s_sk = NKO.simple_kriging(X, y, theta, 2.0, -0.1, "matern5_2", Xn, jitter=NKO.JITTER)
    np.testing.assert_allclose(m_nk, m_sk, rtol=1e-9, atol=1e-9)
    # NK var = sigma2 (1 - k'R^-1 k - jitter): the jitter of KM_gg is the only difference
    np.testing.assert_allclose(s_nk ** 2, np.maximum(s_sk ** 2 - 2.0 * NKO.JITTER, 0), rtol=0, atol=1e-8)


def test_nk_interpolates_and_bounds():
    rng = np.random.default_rng(5)
    X = rng.uniform(size=(100, 2))
    y = f2d(X)
    groups = [np.arange(g, 100, 4) for g in range(4)]
    theta = np.array([0.3, 0.3])
    mean, sd = NKO.predict_nk(X, y, groups, theta, 1.0, 0.2, "matern5_2", X)
    assert np.max(np.abs(mean - y)) < 1e-3
    assert sd.max() < 1e-2
    Xt = rng.uniform(size=(40, 2))
    m2, s2 = NKO.predict_nk(X, y, groups, theta, 1.0, 0.2, "matern5_2", Xt)
    assert np.all(s2 ** 2 <= 1.0 * (1 + 1e-6))
    mf, sf = NKO.predict_nk(X, y, groups, theta, 1.0, 0.2, "matern5_2", np.full((3, 2), 10.0))
    assert np.all(sf ** 2 > 0.5) and np.max(np.abs(mf - 0.2)) < 1e-6


def test_poe_family_single_expert():
    """PoE / gPoE / BCM of one expert are the expert.  rBCM is not: the code
    (NestedKriging.cpp:400-406) adds (1 - beta) / prior_var to the precision,
    so its mean is beta mu / v / (beta / v + (1 - beta) / prior_var), checked
    here in closed form (the reference's test comment expects the mean to be
    unchanged; the code it tests does not do that)."""
    rng = np.random.default_rng(2)
    mus = rng.normal(size=(20, 1))
    sds = rng.uniform(0.1, 0.9, size=(20, 1))
    for agg in ("PoE", "gPoE", "BCM"):
        m, s = NKO.poe_family(mus, sds, agg, 1.0)
        np.testing.assert_allclose(m, mus[:, 0], rtol=1e-12)
        np.testing.assert_allclose(s, np.sqrt(sds[:, 0] ** 2 + 1e-12), rtol=1e-12)
    m, s = NKO.poe_family(mus, sds, "rBCM", 2.0)
    v = sds[:, 0] ** 2 + 1e-12
    beta = 0.5 * (np.log(2.0) - np.log(v))
    prec = beta / v + (1 - beta) / 2.0
    np.testing.assert_allclose(m, beta * mus[:, 0] / v / prec, rtol=1e-12)
    np.testing.assert_allclose(s, np.sqrt(1 / prec), rtol=1e-12)


def test_poe_is_overconfident():
    """PoE precision is the sum of the experts' precisions: its variance is
    below every expert's (NestedKrigingTest.cpp:129-135)"""
    rng = np.random.default_rng(4)
    mus = rng.normal(size=(15, 5))
    sds = rng.uniform(0.1, 0.9, size=(15, 5))
    _, s = NKO.poe_family(mus, sds, "PoE", 1.0)
    assert np.all(s[:, None] <= sds + 1e-9)


def test_unify_weighted_means():
    th = [np.array([0.1, 1.0]), np.array([0.4, 4.0])]
    t, s2, b0 = NKO.unify(th, [1.0, 3.0], [0.0, 2.0], [10, 30])
    np.testing.assert_allclose(t, [0.1 ** 0.25 * 0.4 ** 0.75, 1.0 ** 0.25 * 4.0 ** 0.75], rtol=1e-14)
    assert abs(s2 - 2.5) < 1e-15 and abs(b0 - 1.5) < 1e-15
