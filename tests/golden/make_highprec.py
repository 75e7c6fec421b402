"""50-digit evaluations of the log-likelihood and its gradient on badly
conditioned cases (tests/illcond_cases.py), written to
tests/golden/highprec.json.

The formulas are the reference's (Kriging::_logLikelihood,
src/lib/Kriging.cpp:214-338; gradient KrigingImpl.cpp:855-885): R from the
kernel with unit diagonal (times alpha off the diagonal for the nugget
model), the jittered copy R + J with J the cumulative 1e-10 10^q ladder added
to the diagonal in double precision exactly as the reference does, L its
Cholesky factor, F* = L^-1 F, y* = L^-1 y, beta from F*'F*, E* = L^-1 (y - F
beta), x = L^-T E*, R^-1 from L, and the gradient over the lower triangle of
the un-jittered R.  Everything after the double inputs (X, y, theta, alpha,
the jittered diagonal) is evaluated with mpmath at 50 significant digits:
the kernel values, the factorisation, the solves and the sums.  The jitter
count is the one the CPU oracle's retry ladder chooses.

Each case also records the errors of the reference algorithm evaluated in
FP64 under other valid orderings of its factorisation: LAPACK's dpotrf
(scipy; what Armadillo's chol calls) and blocked right-looking Cholesky with
dtrsm panels of 64 and 32 columns (tests/illcond_cases.py cpu_orderings).

The path of config1's fit: the restated BFGS fit loop (oracle LL + scipy
L-BFGS-B) is run once from the reference's start point with every evaluated
theta recorded; five of them (the start, three iterates, the last) are kept.

Run: python tests/golden/make_highprec.py   (~5 minutes on 8 cores)
"""
import json
import os
import sys
from multiprocessing import Pool

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle as O  # noqa: E402
import illcond_cases as C  # noqa: E402

mp.mp.dps = 50
SQ3, SQ5 = mp.sqrt(3), mp.sqrt(5)


def kern_and_dlog(kernel, dx, theta):
    """k(dx; theta) and d ln k / d theta_k (Covariance.cpp:24-179 restated
    from the kernel definitions: ln k is a sum over coordinates)."""
    lnk = mp.mpf(0)
    g = []
    for dk, tk in zip(dx, theta):
        a = abs(dk)
        if kernel == "gauss":
            lnk -= (dk / tk) ** 2 / 2
            g.append(dk * dk / tk ** 3)
        elif kernel == "exp":
            lnk -= a / tk
            g.append(a / tk ** 2)
        elif kernel == "matern3_2":
            u = SQ3 * a / tk
            lnk += mp.log(1 + u) - u
            g.append(u * u / (1 + u) / tk)
        elif kernel == "matern5_2":
            u = SQ5 * a / tk
            q = 1 + u + u * u / 3
            lnk += mp.log(q) - u
            g.append(u * u * (1 + u) / 3 / q / tk)
        else:
            raise ValueError(kernel)
    return mp.exp(lnk), g


def highprec_ll_grad(X, y, kernel, regmodel, noise, gamma, inc, num_nugget=1e-10):
    n, d = X.shape
    F = O.trend_matrix(X, regmodel)
    p = F.shape[1]
    theta = [mp.mpf(float(t)) for t in gamma[:d]]
    nug = noise == "nugget"
    alpha = mp.mpf(float(gamma[d])) if nug else mp.mpf(1)
    # jittered diagonal, accumulated in double as LinearAlgebra.cpp:88
    dv = 1.0
    for q in range(inc):
        dv += num_nugget * 10.0 ** q
    Xm = [[mp.mpf(float(v)) for v in row] for row in X]
    R = [[mp.mpf(0)] * n for _ in range(n)]
    G = {}
    for i in range(n):
        for j in range(i):
            dx = [Xm[j][k] - Xm[i][k] for k in range(d)]
            r, g = kern_and_dlog(kernel, dx, theta)
            R[i][j] = alpha * r
            G[i, j] = g
    # Cholesky of the jittered matrix (lower, row-oriented)
    L = [[mp.mpf(0)] * n for _ in range(n)]
    for j in range(n):
        s = mp.mpf(dv) - mp.fsum(L[j][k] ** 2 for k in range(j))
        if s <= 0:
            raise ValueError("not positive definite at this jitter")
        L[j][j] = mp.sqrt(s)
        for i in range(j + 1, n):
            L[i][j] = (R[i][j] - mp.fsum(L[i][k] * L[j][k] for k in range(j))) / L[j][j]

    def fwd(b):
        z = []
        for i in range(n):
            z.append((b[i] - mp.fsum(L[i][k] * z[k] for k in range(i))) / L[i][i])
        return z

    def bwd(b):  # L^T z = b
        z = [mp.mpf(0)] * n
        for i in range(n - 1, -1, -1):
            z[i] = (b[i] - mp.fsum(L[k][i] * z[k] for k in range(i + 1, n))) / L[i][i]
        return z

    ym = [mp.mpf(float(v)) for v in y]
    Fs = [fwd([mp.mpf(float(F[i, a])) for i in range(n)]) for a in range(p)]
    ys = fwd(ym)
    Gm = mp.matrix(p, p)
    h = mp.matrix(p, 1)
    for a in range(p):
        h[a] = mp.fsum(Fs[a][i] * ys[i] for i in range(n))
        for c in range(p):
            Gm[a, c] = mp.fsum(Fs[a][i] * Fs[c][i] for i in range(n))
    beta = mp.lu_solve(Gm, h) if p > 0 else []
    r = [ym[i] - mp.fsum(mp.mpf(float(F[i, a])) * beta[a] for a in range(p)) for i in range(n)]
    E = fwd(r)
    sse = mp.fsum(e * e for e in E)
    logdet = mp.fsum(mp.log(L[i][i]) for i in range(n))
    two_pi = 2 * mp.pi
    if nug:
        var = sse / n
        ll = -(n * mp.log(two_pi * var) + 2 * logdet + sse / var) / 2
        s2 = var
    else:
        s2 = sse / n
        ll = -(n * mp.log(two_pi * s2) + 2 * logdet + n) / 2
    # gradient: x = L^-T E*, R^-1 lower triangle from L^-1
    x = bwd(E)
    Linv = [[mp.mpf(0)] * n for _ in range(n)]  # Linv[i][j], i >= j
    for j in range(n):
        for i in range(j, n):
            t = (1 if i == j else 0) - mp.fsum(L[i][k] * Linv[k][j] for k in range(j, i))
            Linv[i][j] = t / L[i][i]
    t1 = [mp.mpf(0)] * d
    t2 = [mp.mpf(0)] * d
    qa = mp.mpf(0)
    tra = mp.mpf(0)
    for i in range(n):
        for j in range(i):
            rinv = mp.fsum(Linv[k][i] * Linv[k][j] for k in range(i, n))
            Rij = R[i][j]
            for k in range(d):
                gk = Rij * G[i, j][k]
                t1[k] += 2 * x[i] * gk * x[j]
                t2[k] -= 2 * rinv * gk
            qa += 2 * x[i] * (Rij / alpha) * x[j]
            tra += 2 * rinv * (Rij / alpha)
    grad = [(t1[k] / s2 + t2[k]) / 2 for k in range(d)]
    if nug and len(gamma) > d:
        grad.append(-(-qa / s2 + tra) / 2)
    return ll, grad


def c1_path():
    from fit_restatement import mt19937_uniform
    from scipy.optimize import fmin_l_bfgs_b
    X, y = C.config1()
    P = O.OracleProblem(X, y, "gauss", "constant")
    lo, up = O.theta_bounds(X, y)
    start = lo + mt19937_uniform(123, 3) * (up - lo)
    seen = []

    def fofn(g):
        th = np.exp(g)
        seen.append(th.copy())
        ll, gr, _ = P.log_likelihood(th)
        return -ll, -(gr * th)

    gamma = np.log(start)
    fmin_l_bfgs_b(fofn, gamma, bounds=list(zip(np.minimum(gamma, np.log(lo)), np.maximum(gamma, np.log(up)))),
                  m=5, factr=1e10, pgtol=1e-3, maxiter=20)
    k = len(seen)
    pick = sorted(set([0, k // 4, k // 2, (3 * k) // 4, k - 1]))
    return np.array([seen[q] for q in pick])


def _one(item):
    name, (X, y, kernel, regmodel, noise, gamma, chol) = item
    P = O.OracleProblem(X, y, kernel, regmodel, noise)
    cfg = O.default_cfg(**chol) if chol else None
    _, _, inc = P.log_likelihood(gamma, cfg=cfg)
    nn = chol["num_nugget"] if chol else 1e-10
    ll, grad = highprec_ll_grad(X, y, kernel, regmodel, noise, gamma, inc, num_nugget=nn)
    # errors of the reference algorithm under the CPU orderings of its
    # factorisation (LAPACK, blocked dtrsm panels), against these values
    llf, gf = float(ll), np.array([float(g) for g in grad])
    spread = {}
    for k, (l_o, g_o) in C.cpu_orderings(X, y, kernel, regmodel, noise, gamma, inc, nn).items():
        spread[k] = {"ll_err": abs(l_o - llf) / abs(llf),
                     "grad_err": float(np.linalg.norm(g_o - gf) / max(np.linalg.norm(gf), 1e-300))}
    return name, {"gamma": [float(v) for v in gamma], "inc": int(inc), "ll": mp.nstr(ll, 40),
                  "grad": [mp.nstr(g, 40) for g in grad], "cpu_orderings": spread}


def main():
    path = c1_path()
    cases = C.cases(path)
    with Pool(8) as pool:
        res = dict(pool.map(_one, sorted(cases.items(), key=lambda kv: -kv[1][0].shape[0]), chunksize=1))
    out = {"dps": 50, "c1_path": path.tolist(), "cases": res}
    with open(os.path.join(HERE, "highprec.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(res), "cases")


if __name__ == "__main__":
    main()
