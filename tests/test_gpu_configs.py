"""The single-evaluation and adjoint configurations at their full sizes,
checked through properties that do not need the (cubic, single-threaded)
oracle: an independent dense numpy/LAPACK formulation of the LL and of the
analytic gradient (W = x x'/sigma2 - R^-1 contracted with dR/dtheta_k,
Kriging.cpp:291-326), and central finite differences of the GPU's own LL.
The dense formulation is held to the plain bars (LL 1e-10 relative,
gradient 1e-8 normwise); the measured errors are printed."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def corr(X, theta, kernel):
    """dense correlation matrix, built column block by column block"""
    n = X.shape[0]
    R = np.empty((n, n))
    for j0 in range(0, n, 1024):
        D = np.abs(X[:, None, :] - X[None, j0:j0 + 1024, :]) / theta
        if kernel == "gauss":
            R[:, j0:j0 + 1024] = np.exp(-0.5 * np.sum(D * D, axis=2))
        elif kernel == "matern5_2":
            v = np.sqrt(5.0) * D
            R[:, j0:j0 + 1024] = np.exp(-np.sum(v, axis=2)) * np.prod(1 + v + v * v / 3.0, axis=2)
        else:
            raise ValueError(kernel)
    return R


def dense_ll_grad(X, y, F, theta, kernel, alpha=None):
    """LL and gradient (theta, and alpha for the nugget model with sigma2 and
    nugget both estimated) by dense LAPACK algebra."""
    n, d = X.shape
    C = corr(X, theta, kernel)
    R = C * (alpha if alpha is not None else 1.0)
    np.fill_diagonal(R, 1.0)
    L = np.linalg.cholesky(R)
    Li = np.linalg.solve(L, np.eye(n))
    Fs, ys = Li @ F, Li @ y
    beta = np.linalg.lstsq(Fs, ys, rcond=None)[0]
    E = Li @ (y - F @ beta)
    s2 = E @ E / n
    logdet = np.sum(np.log(np.diag(L)))
    ll = -0.5 * (n * np.log(2 * np.pi * s2) + 2 * logdet + n)
    x = Li.T @ E
    Rinv = Li.T @ Li
    W = np.outer(x, x) / s2 - Rinv
    np.fill_diagonal(W, 0.0)
    g = np.zeros(d + (1 if alpha is not None else 0))
    for k in range(d):
        dk = X[:, k][:, None] - X[:, k][None, :]
        if kernel == "gauss":
            G = dk * dk / theta[k] ** 3
        else:
            u = np.sqrt(5.0) * np.abs(dk) / theta[k]
            a, b = 1 + u, u * u / 3
            G = a * b / (a + b) / theta[k]
        g[k] = 0.5 * np.sum(W * R * G)
    if alpha is not None:
        g[d] = 0.5 * np.sum(W * C)
    return ll, g, np.linalg.cond(R) if n <= 4096 else None


def fd_grad(fun, gamma, rel):
    g = np.zeros_like(gamma)
    for k in range(len(gamma)):
        h = rel * abs(gamma[k])
        gp, gm = gamma.copy(), gamma.copy()
        gp[k] += h
        gm[k] -= h
        g[k] = (fun(gp) - fun(gm)) / (2 * h)
    return g


def test_single_evaluation_config_n4096():
    """n = 4096, d = 8, matern5_2, linear trend, theta ~ log-U[0.2, 1];
    y = 1 + sum_j x_j + a GP sample drawn with the same kernel (sigma2 = 1)."""
    from libkriging_amd import _native as N
    rng = np.random.default_rng(124)
    n, d = 4096, 8
    X = rng.uniform(size=(n, d))
    theta = np.exp(rng.uniform(np.log(0.2), np.log(1.0), size=d))
    R = corr(X, theta, "matern5_2")
    Lg = np.linalg.cholesky(R + 1e-10 * np.eye(n))
    y = 1.0 + X.sum(axis=1) + Lg @ rng.normal(size=n)
    F = np.hstack([np.ones((n, 1)), X])
    ctx = N.Context(0)
    pid = ctx.problem(X, y, F, "matern5_2")
    ll, g, st, inc = ctx.ll_grad_batch([pid], theta[None, :], True)
    assert st[0] == 0 and inc[0] == 0
    ll_d, g_d, c = dense_ll_grad(X, y, F, theta, "matern5_2")
    e_ll, e_g = abs(ll[0] - ll_d) / abs(ll_d), rel_err(g[0], g_d)
    print(f"n4096: LL {ll[0]!r} dense {ll_d!r} rel err {e_ll:.2e}; grad rel err {e_g:.2e}; cond(R) {c:.2e}")
    assert e_ll <= 1e-10
    assert e_g <= 1e-8
    fd = fd_grad(lambda t: ctx.ll_grad_batch([pid], t[None, :], False)[0][0], theta, 1e-6)
    assert rel_err(g[0], fd) <= 1e-5


def test_adjoint_nugget_config_n8192():
    """n = 8192, d = 10, gauss, theta ~ log-U[0.3, 1.5], nugget ratio 1e-6
    (alpha = 1 / (1 + 1e-6)); y = GP sample + N(0, 1e-3) noise.  Gradient over
    the 10 thetas and alpha from the fused adjoint kernel, behind the model's
    logLikelihoodFun and the torch.autograd.Function, against the dense
    analytic gradient, central finite differences (relative step 1e-6 in
    theta; in gamma_alpha = -log(1 + 1e-3 - alpha) for alpha so the step
    cannot cross alpha = 1), and the layer-1 batch entry point."""
    import torch
    import libkriging_amd as lk
    from libkriging_amd import _native as N
    rng = np.random.default_rng(125)
    n, d = 8192, 10
    X = rng.uniform(size=(n, d))
    theta = np.exp(rng.uniform(np.log(0.3), np.log(1.5), size=d))
    alpha = 1.0 / (1.0 + 1e-6)
    C = corr(X, theta, "gauss")
    Rg = alpha * C
    np.fill_diagonal(Rg, 1.0)
    Lg = np.linalg.cholesky(Rg)
    y = Lg @ rng.normal(size=n) + rng.normal(scale=np.sqrt(1e-3), size=n)
    del C, Rg, Lg
    F = np.ones((n, 1))
    ctx = N.Context(0)
    pid = ctx.problem(X, y, F, "gauss", "nugget")
    gam = np.concatenate([theta, [alpha]])
    ll, g, st, inc = ctx.ll_grad_batch([pid], gam[None, :], True)
    assert st[0] == 0
    # the model surface: logLikelihoodFun and the torch Function give the
    # batch entry point's values bit for bit
    k = lk.Kriging(y, X, "gauss", "constant", False, "none", "LL",
                   {"theta": theta[None, :], "sigma2": 1.0, "nugget": 1e-6}, "nugget")
    ll_k, g_k, _ = k.logLikelihoodFun(gam, True)
    assert ll_k == ll[0] and np.array_equal(g_k, g[0])
    t = torch.tensor(gam, dtype=torch.float64, requires_grad=True)
    v = lk.log_likelihood(t, k)
    v.backward()
    assert v.item() == ll[0] and np.array_equal(t.grad.numpy(), g[0])
    ll_d, g_d, _ = dense_ll_grad(X, y, F, theta, "gauss", alpha)
    e_ll, e_g = abs(ll[0] - ll_d) / abs(ll_d), rel_err(g[0], g_d)
    print(f"n8192: LL {ll[0]!r} dense {ll_d!r} rel err {e_ll:.2e}; grad rel err {e_g:.2e} "
          f"(theta part {rel_err(g[0][:d], g_d[:d]):.2e}, alpha {abs(g[0][d] - g_d[d]) / abs(g_d[d]):.2e}); "
          f"jitter steps {inc[0]}")
    assert e_ll <= 1e-10
    assert e_g <= 1e-8

    def ll_of(gm):
        return ctx.ll_grad_batch([pid], gm[None, :], False)[0][0]

    fd_theta = np.array([(ll_of(np.concatenate([theta + (np.arange(d) == k) * 1e-6 * theta[k], [alpha]])) -
                          ll_of(np.concatenate([theta - (np.arange(d) == k) * 1e-6 * theta[k], [alpha]])))
                         / (2e-6 * theta[k]) for k in range(d)])
    ga =-np.log(1 + 1e-3 - alpha)
    h = 1e-6 * ga

    def ll_ga(v):
        return ll_of(np.concatenate([theta, [1 + 1e-3 - np.exp(-v)]]))

    dll_dga = (ll_ga(ga + h) - ll_ga(ga - h)) / (2 * h)
    fd_alpha = dll_dga / (1 + 1e-3 - alpha)  # d gamma_alpha / d alpha = 1 / (1 + 1e-3 - alpha)
    print("fd theta", fd_theta, "fd alpha", fd_alpha)
    assert rel_err(g[0][:d], fd_theta) <= 1e-4
    assert abs(g[0][d] - fd_alpha) <= 1e-4 * max(abs(fd_alpha), np.linalg.norm(g[0]))


def test_torch_autograd_function():
    import torch
    import libkriging_amd as lk
    rng = np.random.default_rng(3)
    X = rng.uniform(size=(120, 3))
    y = np.prod(1.0 + 0.5 * np.cos(3.0 * X), axis=1)
    k = lk.Kriging(y, X, "matern3_2", "constant", False, "none", "LL", {"theta": np.array([[0.4, 0.5, 0.6]])})
    th = torch.tensor([0.3, 0.45, 0.7], dtype=torch.float64, requires_grad=True)
    ll = lk.log_likelihood(th, k)
    ll.backward()
    ll_r, g_r, _ = k.logLikelihoodFun(th.detach().numpy(), True)
    assert abs(ll.item() - ll_r) <= 1e-14 * abs(ll_r)
    assert np.allclose(th.grad.numpy(), g_r, rtol=1e-14, atol=0)
    # batched: rows are independent, backward of the sum gives each row's gradient
    thb = torch.tensor(np.exp(rng.uniform(-1.5, 0.0, size=(5, 3))), requires_grad=True)
    lk.log_likelihood(thb, k).sum().backward()
    for i in range(5):
        _, gi, _ = k.logLikelihoodFun(thb.detach().numpy()[i], True)
        assert np.allclose(thb.grad.numpy()[i], gi, rtol=1e-13)
    # against finite differences through torch's own checker
    assert torch.autograd.gradcheck(lambda t: lk.log_likelihood(t, k), (th.detach().clone().requires_grad_(),),
                                    eps=1e-6, atol=1e-6, rtol=1e-5)
