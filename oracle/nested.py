"""CPU restatement of NestedKriging's unification and aggregation steps.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker; the product
package never imports it).  numpy / scipy LAPACK for the small dense algebra,
the C oracle (lk_oracle.c) for the kernel values.

Follows src/lib/NestedKriging.cpp:
  unify            unify_hyperparameters, plain path (:308-338)
  precompute_nk    precompute_nk (:340-351)
  predict_nk       predict_nk (:353-465), same chunked loop structure
  poe_family       predict_poe_family (:365-421) from submodel (mean, stdev)
Pins (tests/test_nested_oracle.py): singleton groups reduce the NK predictor
to full simple kriging (an independent dense solve); one group reduces it to
that group's simple-kriging predictor; interpolation of the design; PoE-family
closed forms at p = 1; the variance bound var <= sigma2 and the far-field
limit (mean -> beta0, var -> sigma2).
"""
import numpy as np
import scipy.linalg as sla

from . import oracle as O

JITTER = 1e-10  # NestedKriging::jitter (NestedKriging.hpp:165)


def unify(thetas, sigma2s, beta0s, sizes):
    """weighted geometric mean of theta, weighted sigma2 and beta0, weights n_g / n"""
    sizes = np.asarray(sizes, dtype=np.float64)
    n = sizes.sum()
    log_theta = np.zeros(len(thetas[0]))
    s2 = 0.0
    b0 = 0.0
    for g in range(len(sizes)):
        w = sizes[g] / n
        log_theta += w * np.log(np.asarray(thetas[g]))
        s2 += w * sigma2s[g]
        if beta0s is not None:
            b0 += w * beta0s[g]
    return np.exp(log_theta), s2, b0


def precompute_nk(X, y, groups, theta, beta0, kernel, jitter=JITTER):
    """per group: L_g = chol(R_g + jitter I), alpha_g = R_g^-1 (y_g - beta0)"""
    Ls, alphas = [], []
    for idx in groups:
        Xg = X[idx]
        R = O.corr_mat(kernel, Xg, Xg, theta)
        R[np.diag_indices_from(R)] += jitter
        L = np.linalg.cholesky(R)
        r = y[idx] - beta0
        a = sla.solve_triangular(L.T, sla.solve_triangular(L, r, lower=True), lower=False)
        Ls.append(L)
        alphas.append(a)
    return Ls, alphas


def _solve_km(KM, kM):
    """arma::solve(symmatu(KM), kM, likely_sympd), pinv on failure"""
    try:
        c = sla.cho_factor(KM, lower=False, check_finite=True)
        return sla.cho_solve(c, kM), 0
    except np.linalg.LinAlgError:
        pass
    try:
        return np.linalg.solve(KM, kM), 1
    except np.linalg.LinAlgError:
        return np.linalg.pinv(KM) @ kM, 2


_PAIRS = None  # (X, groups, W, theta, kernel) for the pair workers


def _pair_cross(ij):
    X, groups, W, theta, kernel = _PAIRS
    i, j = ij
    Rij = O.corr_mat(kernel, X[groups[i]], X[groups[j]], theta)
    return np.sum(W[i] * (Rij @ W[j]), axis=0)


def predict_nk(X, y, groups, theta, sigma2, beta0, kernel, Xn, chunk=128, return_info=False, jitter=JITTER,
               workers=1):
    """jitter: the reference's 1e-10; other values only for the pins.
    workers > 1: the group-pair sums of a chunk over that many forked
    processes (same arithmetic per pair)"""
    global _PAIRS
    X = np.asarray(X, dtype=np.float64)
    Xn = np.asarray(Xn, dtype=np.float64)
    Ls, alphas = precompute_nk(X, np.asarray(y, dtype=np.float64), groups, theta, beta0, kernel, jitter)
    p = len(groups)
    q = Xn.shape[0]
    mean = np.empty(q)
    var = np.empty(q)
    info = np.zeros(q, dtype=int)
    for start in range(0, q, chunk):
        Xc = Xn[start:start + chunk]
        qc = Xc.shape[0]
        W = []
        M = np.empty((p, qc))
        Kd = np.empty((p, qc))
        for g, idx in enumerate(groups):
            C = O.corr_mat(kernel, X[idx], Xc, theta)
            U = sla.solve_triangular(Ls[g], C, lower=True)
            W.append(sla.solve_triangular(Ls[g].T, U, lower=False))
            M[g] = beta0 + C.T @ alphas[g]
            Kd[g] = np.sum(U * U, axis=0)
        cross = np.zeros((p, p, qc))
        pairs = [(i, j) for i in range(p) for j in range(i + 1, p)]
        _PAIRS = (X, groups, W, theta, kernel)
        if workers > 1:
            import multiprocessing as mproc
            with mproc.get_context("fork").Pool(workers) as pool:
                sums = pool.map(_pair_cross, pairs, chunksize=max(1, len(pairs) // (16 * workers)))
        else:
            sums = [_pair_cross(ij) for ij in pairs]
        _PAIRS = None
        for (i, j), cij in zip(pairs, sums):
            cross[i, j] = cij
            cross[j, i] = cij
        for t in range(qc):
            KM = cross[:, :, t].copy()
            KM[np.diag_indices(p)] = Kd[:, t] + jitter
            kM = np.diag(KM).copy()
            w, inf = _solve_km(np.triu(KM) + np.triu(KM, 1).T, kM)
            info[start + t] = inf
            mean[start + t] = beta0 + w @ (M[:, t] - beta0)
            var[start + t] = sigma2 * max(0.0, 1.0 - w @ kM)
    if return_info:
        return mean, np.sqrt(var), info
    return mean, np.sqrt(var)


def poe_family(mus, sds, aggregation, prior_var):
    """mus, sds: (q, p) submodel means / standard deviations"""
    tiny = 1e-12
    mus = np.asarray(mus, dtype=np.float64)
    vars_ = np.asarray(sds, dtype=np.float64) ** 2 + tiny
    p = mus.shape[1]
    if aggregation == "PoE":
        prec = np.sum(1.0 / vars_, axis=1)
        mean = np.sum(mus / vars_, axis=1) / prec
    elif aggregation == "gPoE":
        prec = np.sum(1.0 / vars_, axis=1) / p
        mean = (np.sum(mus / vars_, axis=1) / p) / prec
    elif aggregation == "BCM":
        prec = np.sum(1.0 / vars_, axis=1) - (p - 1.0) / prior_var
        prec = np.clip(prec, tiny, np.inf)
        mean = np.sum(mus / vars_, axis=1) / prec
    elif aggregation == "rBCM":
        beta = 0.5 * (np.log(prior_var) - np.log(vars_))
        prec = np.sum(beta / vars_, axis=1) + (1.0 - np.sum(beta, axis=1)) / prior_var
        prec = np.clip(prec, tiny, np.inf)
        mean = np.sum(beta * mus / vars_, axis=1) / prec
    else:
        raise ValueError(aggregation)
    return mean, np.sqrt(np.clip(1.0 / prec, 0.0, np.inf))


def simple_kriging(X, y, theta, sigma2, beta0, kernel, Xn, jitter=0.0):
    """full-design simple kriging (known beta0, sigma2), by a dense solve:
    mean = beta0 + r' R^-1 (y - beta0), var = sigma2 (1 - r' R^-1 r)"""
    R = O.corr_mat(kernel, X, X, theta)
    R[np.diag_indices_from(R)] += jitter
    r = O.corr_mat(kernel, X, Xn, theta)
    A = np.linalg.solve(R, np.column_stack([np.asarray(y) - beta0, r]))
    mean = beta0 + r.T @ A[:, 0]
    var = sigma2 * np.maximum(0.0, 1.0 - np.sum(r * A[:, 1:], axis=0))
    return mean, np.sqrt(var)
