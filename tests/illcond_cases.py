"""Badly conditioned evaluations whose GPU and oracle results are both
compared with a 50-digit evaluation of the same formulas
(tests/golden/make_highprec.py writes it to tests/golden/highprec.json).

Each case: (X, y, kernel, regmodel, noise, gamma, chol), chol None (the
reference's defaults) or the retry-ladder settings of the case (num_nugget,
rcond_check, max_inc, min_rcond).  The jitter count used by the
high-precision evaluation is the one the oracle's retry ladder chose under
those settings; the GPU tests check that the GPU chose the same.
"""
import numpy as np


def _design(seed, n, d):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(2 * np.pi * X).sum(axis=1) + 0.5 * X[:, 0] * X[:, -1]
    return X, y


def near_duplicate():
    rng = np.random.default_rng(77)
    X = rng.uniform(size=(30, 2))
    X[29] = X[0] + 3e-7
    return X, np.sin(3 * X).sum(axis=1)


def config1(seed=124, n=200, d=3):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(2 * np.pi * X).sum(axis=1) + 0.5 * X[:, 0] * X[:, 1]
    return X, y


# five parameter vectors on the path of the restated BFGS fit of config1
# (gauss kernel, n = 200, d = 3): the start, three iterates and the optimum
# (written by make_highprec.py into the fixture; listed here after a run)
def cases(c1_path=None):
    out = {}
    # duplicated design rows: R singular, the retry ladder adds jitter
    X, y = _design(9, 80, 2)
    X = np.vstack([X, X[:5]])
    y = np.concatenate([y, y[:5]])
    out["duplicates"] = (X, y, "gauss", "constant", None, np.array([0.3, 0.3]), None)
    # the gauss designs of the kernel sweep whose R is badly conditioned
    # (cond(R) 1e7 .. 1e14): 37 points of [0,1]^2, 64 of [0,1]^3, 300 of
    # [0,1]^4, theta log-uniform in [0.2, 1]
    for n, d, slots in ((37, 2, (0, 1, 2)), (64, 3, (0, 1, 2)), (300, 4, (0,))):
        X, y = _design(n * 7 + d, n, d)
        gam = np.exp(np.random.default_rng(n + d).uniform(np.log(0.2), np.log(1.0), size=(3, d)))
        for b in slots:
            out[f"gauss{n}_{b}"] = (X, y, "gauss", "constant", None, gam[b], None)
    # the n = 8192, d = 10 gauss + nugget configuration on 256- and 384-point
    # subsamples (nugget ratio 1e-6: alpha = 1 / (1 + 1e-6))
    rng = np.random.default_rng(125)
    Xf = rng.uniform(size=(1024, 10))
    th = np.exp(rng.uniform(np.log(0.3), np.log(1.5), size=10))
    yf = np.sin(3 * Xf).sum(axis=1) + rng.normal(0.0, np.sqrt(1e-3), size=1024)
    alpha = 1.0 / (1.0 + 1e-6)
    for m in (256, 384):
        out[f"nugget_d10_n{m}"] = (Xf[:m], yf[:m], "gauss", "constant", "nugget", np.concatenate([th, [alpha]]), None)
    # a near-duplicate design point (3e-7 away): R factors, but rcond(L)^2 is
    # 8.5e-14; with the condition threshold at 1e-12 the check rejects the
    # factor and one jitter step is added (inc 1), without the check it is
    # kept (inc 0)
    X, y = near_duplicate()
    for check in (1, 0):
        out[f"near_duplicate_check{check}"] = (X, y, "matern5_2", "constant", None, np.array([0.2, 0.2]),
                                               dict(num_nugget=1e-10, rcond_check=check, max_inc=10, min_rcond=1e-12))
    if c1_path is not None:
        X, y = config1()
        for q, g in enumerate(np.atleast_2d(c1_path)):
            out[f"c1_path_{q}"] = (X, y, "gauss", "constant", None, np.asarray(g), None)
    return out


LL_BAR, GRAD_BAR = 1e-10, 1e-8


def load_truth():
    """(cases, truth) from tests/golden/highprec.json"""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "highprec.json")
    with open(path) as f:
        hp = json.load(f)
    return cases(np.array(hp["c1_path"])), hp["cases"]


def arbitrate(truth, ll_gpu, g_gpu, ll_orc, g_orc):
    """Errors against the 50-digit values and whether the GPU passes.  The
    reference's own FP64 arithmetic is one of several valid orderings of its
    factorisation (the C oracle's unblocked one; LAPACK's dpotrf, which a
    reference build calls through Armadillo; blocked dtrsm-panel variants);
    on these matrices their errors differ by up to two orders of magnitude,
    so the GPU must stay within max(plain bar, 2 x the largest error among
    them): never less accurate than the reference's algorithm in FP64."""
    ll_t = float(truth["ll"])
    g_t = np.array([float(v) for v in truth["grad"]])
    gn = max(np.linalg.norm(g_t), 1e-300)
    e = dict(ll_gpu=abs(ll_gpu - ll_t) / abs(ll_t), ll_orc=abs(ll_orc - ll_t) / abs(ll_t),
             g_gpu=np.linalg.norm(np.asarray(g_gpu) - g_t) / gn, g_orc=np.linalg.norm(np.asarray(g_orc) - g_t) / gn)
    sp = truth.get("cpu_orderings", {})
    e["ll_cpu_max"] = max([e["ll_orc"]] + [v["ll_err"] for v in sp.values()])
    e["g_cpu_max"] = max([e["g_orc"]] + [v["grad_err"] for v in sp.values()])
    ok = e["ll_gpu"] <= max(LL_BAR, 2 * e["ll_cpu_max"]) and e["g_gpu"] <= max(GRAD_BAR, 2 * e["g_cpu_max"])
    return e, ok


def _dlnk(kernel, D, th):
    """d ln k / d theta of one coordinate for all pairs: D = x_i - x_j (matrix)"""
    a = np.abs(D)
    if kernel == "gauss":
        return D * D / th ** 3
    if kernel == "exp":
        return a / th ** 2
    if kernel == "matern3_2":
        u = np.sqrt(3.0) * a / th
        return u * u / (1 + u) / th
    u = np.sqrt(5.0) * a / th
    return u * u * (1 + u) / 3 / (1 + u + u * u / 3) / th


def ll_grad_from_factor(X, y, kernel, regmodel, noise, gamma, R, L):
    """The reference's LL and gradient (Kriging.cpp:214-338,
    KrigingImpl.cpp:855-885) from a given lower factor L of the jittered R
    (R itself un-jittered, for the gradient), with LAPACK triangular solves
    (dtrtrs, as arma::solve(trimatl / trimatu)) and R^-1 = L^-T (L^-1 I)
    (inv_sympd).  Sigma2 (and the nugget) estimated."""
    import scipy.linalg as sla
    from oracle import oracle as O
    n, d = X.shape
    F = O.trend_matrix(X, regmodel)
    lo = lambda b: sla.solve_triangular(L, b, lower=True)
    Fs, ys = lo(F), lo(y)
    if F.shape[1] > 0:
        Rs = np.linalg.cholesky(Fs.T @ Fs).T
        beta = sla.solve_triangular(Rs, sla.solve_triangular(Rs.T, Fs.T @ ys, lower=True), lower=False)
    else:
        beta = np.zeros(0)
    E = lo(y - F @ beta)
    sse = E @ E
    logdet = np.sum(np.log(np.diag(L)))
    s2 = sse / n
    ll = -0.5 * (n * np.log(2 * np.pi * s2) + 2 * logdet + n)
    x = sla.solve_triangular(L.T, E, lower=False)
    Rinv = sla.solve_triangular(L.T, lo(np.eye(n)), lower=False)
    low = np.tril(np.ones((n, n), dtype=bool), -1)
    W = (np.outer(x, x) / s2 - Rinv) * R
    g = [np.sum((W * _dlnk(kernel, X[:, k][:, None] - X[:, k][None, :], gamma[k]))[low]) for k in range(d)]
    if noise == "nugget" and len(gamma) > d:
        alpha = gamma[d]
        M = (np.outer(x, x) / s2 - Rinv) * (R / alpha)
        g.append(np.sum(M[low]))  # -(1/2)(-q/s2 + tr) over i != j, both triangles
    return ll, np.array(g)


def jittered_R(X, y, kernel, regmodel, noise, gamma, inc, num_nugget=1e-10):
    from oracle import oracle as O
    d = X.shape[1]
    P = O.OracleProblem(X, y, kernel, regmodel, noise)
    R = P.build_R(gamma[:d], gamma[d] if noise == "nugget" else 1.0)
    np.fill_diagonal(R, 1.0)
    dv = 1.0
    for q in range(inc):
        dv += num_nugget * 10.0 ** q
    Rj = R.copy()
    np.fill_diagonal(Rj, dv)
    return R, Rj


def blocked_cholesky(A, nb):
    """right-looking blocked Cholesky with substitution (dtrsm) panels"""
    import scipy.linalg as sla
    A = A.copy()
    n = A.shape[0]
    L = np.zeros_like(A)
    for k0 in range(0, n, nb):
        k1 = min(k0 + nb, n)
        Lkk = np.linalg.cholesky(A[k0:k1, k0:k1])
        L[k0:k1, k0:k1] = Lkk
        if k1 < n:
            P = sla.solve_triangular(Lkk, A[k1:, k0:k1].T, lower=True).T
            L[k1:, k0:k1] = P
            A[k1:, k1:] -= P @ P.T
    return L


def cpu_orderings(X, y, kernel, regmodel, noise, gamma, inc, num_nugget=1e-10):
    """(ll, grad) of the reference's algorithm under valid FP64 orderings of
    its factorisation: LAPACK's dpotrf (what Armadillo's chol calls; here
    scipy's OpenBLAS build) and blocked right-looking Cholesky with dtrsm
    panels of 64 and 32 columns.  Which ordering a reference build uses
    depends on the LAPACK it links; on matrices with cond(R) ~ 1e17 their
    errors differ by up to two orders of magnitude."""
    R, Rj = jittered_R(X, y, kernel, regmodel, noise, gamma, inc, num_nugget)
    out = {}
    for name, L in (("lapack", np.linalg.cholesky(Rj)), ("blocked64", blocked_cholesky(Rj, 64)),
                    ("blocked32", blocked_cholesky(Rj, 32))):
        out[name] = ll_grad_from_factor(X, y, kernel, regmodel, noise, gamma, R, L)
    return out
