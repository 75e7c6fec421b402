"""CPU restatement of the leave-one-out objective (Kriging::_leaveOneOut,
src/lib/Kriging.cpp:353-468) and of the log marginal posterior
(Kriging::_logMargPost, Kriging.cpp:488-650) for small n, in numpy on top of
the C oracle's covariance functions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
package (libkriging_amd) never imports it.

Pinned by (tests/test_loo_oracle.py):
  - the reference's golden vectors data1-scal-result-leaveOneOut(.Grad).csv
    (tests/references, gauss kernel, constant trend) at 1e-12;
  - brute force: the errors and variances equal those of n universal-kriging
    fits on n - 1 points each (Dubrule's identity), for every kernel and
    trend;
  - central finite differences of the value for every gradient component.
The log marginal posterior has no reference fixture: its value is pinned by
the slogdet closed form of the marginal likelihood and the y-scaling
identity, its gradient by finite differences; agreement with the
reference's own numbers is parity unpinned.
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


def log_marg_post(X, y, kernel, theta, regmodel="constant", alpha=None, est_sigma2=True, sigma2=None,
                  est_nugget=True, nugget=None, want_grad=True):
    """(lmp, grad) following Kriging::_logMargPost (Kriging.cpp:488-650) and
    KrigingImpl::compute_lmp_theta_ans (KrigingImpl.cpp:887-921).  alpha is
    None for the noise-free model, else the nugget model's alpha (R =
    alpha C + (1 - alpha) I) and grad has d + 1 entries.  sigma2 / nugget are
    the model's fixed values (used when not estimated)."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    theta = np.asarray(theta, dtype=np.float64).reshape(-1)
    n, d = X.shape
    nug = alpha is not None
    a_ = float(alpha) if nug else 1.0
    C = O.corr_mat(kernel, X, X, theta)
    R = a_ * C
    np.fill_diagonal(R, 1.0)
    L, _, _ = O.safe_chol_lower(R)
    F = O.trend_matrix(X, regmodel)
    p = F.shape[1]
    Fs = np.linalg.solve(L, F)
    ys = np.linalg.solve(L, y)
    Rinv_X = np.linalg.solve(L.T, Fs)
    if p > 0:
        XtRiX = F.T @ Rinv_X
        LX, _, _ = O.safe_chol_lower(XtRiX)
        P = Rinv_X @ np.linalg.solve(LX.T, np.linalg.solve(LX, Rinv_X.T))
        logdet_lx = np.sum(np.log(np.diag(LX)))
    else:
        P = np.zeros((n, n))
        logdet_lx = 0.0
    yt_Rinv = np.linalg.solve(L.T, ys)
    S_2 = yt_Rinv @ y - y @ P @ y
    if nug:
        if est_sigma2 and est_nugget:
            s2 = S_2 / (n - p)
        elif est_sigma2 or est_nugget:
            s2 = sigma2 / a_
        else:
            s2 = sigma2 + nugget
    elif est_sigma2:
        s2 = S_2 / (n - p)
    else:
        s2 = sigma2
    log_S_2 = np.log(s2 * (n - p))
    lml = -np.sum(np.log(np.diag(L))) - logdet_lx - (n - p) / 2.0 * log_S_2
    a = 0.2
    b = 1.0 / n ** (1.0 / d) * (a + d)
    CL = (X.max(axis=0) - X.min(axis=0)) / n ** (1.0 / d)
    t = np.sum(CL / theta) + ((1.0 - a_) / a_ if nug else 0.0)
    lmp = float(lml + (-b * t + a * np.log(t)))
    if not want_grad:
        return lmp, None
    g = np.zeros(d + (1 if nug else 0))
    if est_sigma2:
        Q_output = yt_Rinv - P @ y
        for k in range(d):
            G = np.zeros((n, n))
            for i in range(n):
                for j in range(i):
                    v = R[i, j] * O.dlncov_dtheta(kernel, X[i] - X[j], theta)[k]
                    G[i, j] = G[j, i] = v
            Wb = np.linalg.solve(L.T, np.linalg.solve(L, G)).T - G @ P
            ans = -np.trace(Wb) / 2.0 + y @ Wb.T @ Q_output / (2.0 * s2)
            g[k] = ans - (a * CL[k] / t - b * CL[k]) / theta[k] ** 2
        if nug:
            if est_sigma2 or est_nugget:
                Gd = R / a_
                np.fill_diagonal(Gd, 0.0)
                Wb = np.linalg.solve(L.T, np.linalg.solve(L, Gd)).T - Gd @ P
                ans = -np.trace(Wb) / 2.0 + y @ Wb.T @ Q_output / (2.0 * s2)
                g[d] = ans - (a / t - b) / a_ ** 2
    else:
        eps = 1e-6
        for k in range(d):
            te = theta.copy()
            te[k] += eps
            g[k] = (log_marg_post(X, y, kernel, te, regmodel, alpha, est_sigma2, sigma2, est_nugget, nugget,
                                  want_grad=False)[0] - lmp) / eps
    return lmp, g
