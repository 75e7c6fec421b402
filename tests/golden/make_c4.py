"""Fixtures of the 64-start configuration (n = 1024, d = 6, exp kernel,
constant trend; X, y and the 64 start rows from bench.workload(), seed 126):

* c4_ll_grad.npz -- the CPU oracle's LL and gradient at the 64 start rows;
* c4_fit.npz     -- the restated BFGS fit loop (oracle LL + scipy's L-BFGS-B,
  tests/fit_restatement.py) from those 64 rows: per-start objective and
  theta, and the best one.

The oracle needs ~0.6 s per evaluation here, so the fit loop (~2 500
evaluations) runs over 8 worker processes in this container; the GPU tests
only read the results.  Run: python tests/golden/make_c4.py
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import bench  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _problem():
    X, y, theta = bench.workload()
    return X, y, theta, O.OracleProblem(X, y, "exp", "constant")


def _ll_grad(k):
    _, _, theta, P = _problem()
    ll, g, inc = P.log_likelihood(theta[k])
    return ll, g, inc


def _fit(k):
    from fit_restatement import fit_one_start
    X, y, theta, P = _problem()
    lo, up = O.theta_bounds(X, y)
    f, th = fit_one_start(P, theta[k], lo, up)
    return f, th


def main():
    X, y, theta, _ = _problem()
    with Pool(8) as pool:
        r = pool.map(_ll_grad, range(theta.shape[0]))
        np.savez(os.path.join(HERE, "c4_ll_grad.npz"), theta=theta, ll=np.array([a[0] for a in r]),
                 grad=np.array([a[1] for a in r]), inc=np.array([a[2] for a in r]))
        print("ll/grad done")
        r = pool.map(_fit, range(theta.shape[0]))
    f = np.array([a[0] for a in r])
    th = np.array([a[1] for a in r])
    k = int(np.argmin(f))
    np.savez(os.path.join(HERE, "c4_fit.npz"), start=theta, objective=f, theta=th, best=k, best_ll=-f[k],
             best_theta=th[k])
    print("best start", k, "ll", -f[k], "theta", th[k])


if __name__ == "__main__":
    main()
