"""The 64-start configuration at its full size (n = 1024, d = 6, exp kernel,
constant trend; X, y and the 64 start rows from bench.workload(), seed 126):

* all 64 LL + gradient evaluations of one batched launch against the CPU
  oracle at the plain bars (LL 1e-10 relative, gradient 1e-8 normwise);
* the 64-start BFGS fit (the reference's "BFGS64" with the 64 given rows,
  Kriging.cpp:1591-2215): batched and one-start-at-a-time runs agree bit for
  bit, and the best LL equals that of the restated fit loop (oracle LL +
  scipy L-BFGS-B) within 1e-8.

The oracle results are fixtures (tests/golden/make_c4.py), since the oracle
needs ~0.6 s per evaluation at this size.
"""
import os

import numpy as np
import pytest

import bench
from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu


def _fixture(name):
    return np.load(os.path.join(GOLDEN, name))


def test_c4_batched_ll_grad_matches_oracle():
    from libkriging_amd import _native as N
    X, y, theta = bench.workload()
    z = _fixture("c4_ll_grad.npz")
    assert np.array_equal(z["theta"], theta)
    ctx = N.Context(0)
    try:
        pid = ctx.problem(X, y, np.ones((X.shape[0], 1)), "exp")
        ll, g, st, inc = ctx.ll_grad_batch([pid] * theta.shape[0], theta, True)
    finally:
        ctx.close()
    assert (st == 0).all()
    assert np.array_equal(inc, z["inc"])
    e_ll = np.abs(ll - z["ll"]) / np.abs(z["ll"])
    e_g = np.array([rel_err(g[b], z["grad"][b]) for b in range(theta.shape[0])])
    print(f"64 evaluations: max LL rel err {e_ll.max():.2e}, max grad rel err {e_g.max():.2e}")
    assert e_ll.max() <= 1e-10
    assert e_g.max() <= 1e-8


def test_c4_bfgs64_fit():
    import libkriging_amd as lk
    X, y, theta = bench.workload()
    z = _fixture("c4_fit.npz")
    k = lk.Kriging(y, X, "exp", "constant", False, "BFGS64", "LL", {"theta": theta})
    k1 = lk.Kriging("exp")
    k1.set_batched_fit(False)
    k1.fit(y, X, "constant", False, "BFGS64", "LL", {"theta": theta})
    ll, ll1 = k.logLikelihood(), k1.logLikelihood()
    print(f"batched fit: LL {ll!r} theta {k.theta()}; one start at a time: LL {ll1!r}; "
          f"restated loop: LL {float(z['best_ll'])!r} theta {z['best_theta']}")
    print("fit statistics (evaluations, batches):", k.fit_statistics(), k1.fit_statistics())
    assert ll == ll1
    assert np.array_equal(k.theta(), k1.theta())
    assert k.sigma2() == k1.sigma2()
    assert abs(ll - z["best_ll"]) <= 1e-8 * abs(z["best_ll"])
    # the committed optimum is one of the optima the restated loop reached
    assert min(rel_err(k.theta(), t) for t in z["theta"]) <= 1e-4
