"""CPU restatement of the leave-one-out objective (Kriging::_leaveOneOut,
src/lib/Kriging.cpp:353-468) for small n, in numpy on top of the C oracle's
covariance functions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
package (libkriging_amd) never imports it.

Pinned by (tests/test_loo_oracle.py):
  - the reference's golden vectors data1-scal-result-leaveOneOut(.Grad).csv
    (tests/references, gauss kernel, constant trend) at 1e-12;
  - brute force: the errors and variances equal those of n universal-kriging
    fits on n - 1 points each (Dubrule's identity), for every kernel and
    trend;
  - central finite differences of the value for every gradient component.
"""
import numpy as np

from oracle import oracle as O


def _corr(kernel, X, theta):
    R = O.corr_mat(kernel, X, X, theta)
    np.fill_diagonal(R, 1.0)
    return R


def leave_one_out(X, y, kernel, theta, regmodel="constant", want_grad=True):
    """(loo, grad, yhat, sigma2loo) following Kriging.cpp:353-468:
    L = safe_chol(R) (KrigingImpl::populate_Model, KrigingImpl.cpp:73-125),
    F* = L^-1 F, y* = L^-1 y, R* = chol_upper(F*'F*), GLS beta, E* = L^-1 (y - F beta),
    By = L^-T E*, Q* = qr_econ(F*).Q, A = Q*' L^-1, B = L^-T L^-1 - A'A,
    sigma2LOO = 1 / diag(B), errors = sigma2LOO * By, loo = |errors|^2 / n;
    gradient (Kriging.cpp:440-466): per k, gradR_k (R_ij dlnk/dtheta_k off
    the diagonal), diagdB = -diagABA(B, gradR_k) (LinearAlgebra.cpp:428-433),
    dsigma2 = -sigma2LOO^2 diagdB, derrors = dsigma2 By - sigma2LOO (B (gradR_k By)),
    grad_k = 2 errors . derrors / n."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    theta = np.asarray(theta, dtype=np.float64).reshape(-1)
    n, d = X.shape
    R = _corr(kernel, X, theta)
    L, _, _ = O.safe_chol_lower(R)
    F = O.trend_matrix(X, regmodel)
    Linv = np.linalg.solve(L, np.eye(n))
    Fs = Linv @ F
    ys = Linv @ y
    p = F.shape[1]
    if p > 0:
        Rs = np.linalg.cholesky(Fs.T @ Fs).T
        beta = np.linalg.solve(Rs, np.linalg.solve(Rs.T, Fs.T @ ys))
    else:
        beta = np.zeros(0)
    Es = Linv @ (y - F @ beta)
    By = Linv.T @ Es
    if p > 0:
        Qs, _ = np.linalg.qr(Fs)
        A = Qs.T @ Linv
        B = Linv.T @ Linv - A.T @ A
    else:
        B = Linv.T @ Linv
    s2 = 1.0 / np.diag(B)
    err = s2 * By
    loo = float(err @ err / n)
    yhat = y - err
    grad = None
    if want_grad:
        grad = np.zeros(d)
        Rm = R.copy()
        for k in range(d):
            G = np.zeros((n, n))
            for i in range(n):
                for j in range(i):
                    g = O.dlncov_dtheta(kernel, X[i] - X[j], theta)[k]
                    G[i, j] = G[j, i] = Rm[i, j] * g
            D = np.triu(2.0 * G)
            np.fill_diagonal(D, np.diag(G))
            diagdB = -np.sum((B @ D) * B, axis=1)
            ds2 = -s2 * s2 * diagdB
            derr = ds2 * By - s2 * (B @ (G @ By))
            grad[k] = 2.0 * err @ derr / n
    return loo, grad, yhat, s2


def leave_one_out_bruteforce(X, y, kernel, theta, regmodel="constant"):
    """errors y_i - yhat_{-i}(x_i) and unit-variance UK variances from n fits
    on n - 1 points each: an independent check of the closed forms above"""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    n = X.shape[0]
    F = O.trend_matrix(X, regmodel)
    err = np.zeros(n)
    var = np.zeros(n)
    for i in range(n):
        keep = np.arange(n) != i
        Xk, yk, Fk = X[keep], y[keep], F[keep]
        Rk = _corr(kernel, Xk, theta)
        r = O.corr_mat(kernel, Xk, X[i:i + 1], theta)[:, 0]
        Ri_r = np.linalg.solve(Rk, r)
        if F.shape[1] > 0:
            Ri_F = np.linalg.solve(Rk, Fk)
            beta = np.linalg.solve(Fk.T @ Ri_F, Ri_F.T @ yk)
            u = F[i] - Fk.T @ Ri_r
            var[i] = 1.0 - r @ Ri_r + u @ np.linalg.solve(Fk.T @ Ri_F, u)
        else:
            beta = np.zeros(0)
            var[i] = 1.0 - r @ Ri_r
        pred = F[i] @ beta + Ri_r @ (yk - Fk @ beta)
        err[i] = y[i] - pred
    return err, var
