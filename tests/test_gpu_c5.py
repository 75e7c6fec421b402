"""The uneven NestedKriging configuration at full size: N = 200 000 points in
[0,1]^5, 256 groups of 300..1500 points (lognormal sizes), matern3_2,
constant trend, then the NK aggregated predictor at the 10 000 prediction
points (bench.nested_workload(), seed 127; NestedKriging.cpp:124-498).

* fit: all 256 groups' BFGS fits in shared batched pipelines; the unified
  prior equals oracle/nested.py's unify of the groups' own fits; groups
  fitted alone (smallest, largest, two others) reach the same optimum as in
  the batch;
* predict at 10 000 points: finite, var <= sigma2, and the design is
  interpolated at sampled design points (NestedKrigingTest.cpp:63-75 bars);
* NK predict with the prior fixed at 8 of the points against the CPU
  oracle's NK predictor (fixture from tests/golden/make_c5.py: the oracle
  needs N^2 / 2 kernel values, minutes on 8 cores).
"""
import os
import time

import numpy as np
import pytest

import bench
from conftest import GOLDEN, rel_err
from oracle import nested as NKO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload():
    return bench.nested_workload()


def test_c5_fit_and_predict(workload):
    import libkriging_amd as lk
    X, y, group_of, sizes, Xn = workload
    nk = lk.NestedKriging("matern3_2")
    t0 = time.perf_counter()
    nk.fit_groups(y, X, group_of, "constant", "BFGS")
    t_fit = time.perf_counter() - t0
    ev, nb = nk.fit_statistics()
    print(f"fit: {t_fit:.2f} s, {ev} LL+grad evaluations in {nb} batched pipelines")
    assert [len(g) for g in nk.groups()] == list(sizes)
    # unification (NestedKriging.cpp:308-322) of the groups' own optima
    th_g, s2_g, b0_g = nk.group_fits()
    t, s2, b0 = NKO.unify(list(th_g), s2_g, b0_g, sizes)
    np.testing.assert_allclose(nk.theta(), t, rtol=1e-12)
    assert abs(nk.sigma2() - s2) <= 1e-12 * s2
    assert abs(nk.beta0() - b0) <= 1e-12 * max(1.0, abs(b0))
    # a group fitted on its own reaches what it reached inside the batch
    order = np.argsort(sizes, kind="stable")
    for g in (order[0], order[-1], order[85], order[170]):
        idx = np.flatnonzero(group_of == g)
        km = lk.Kriging(y[idx], X[idx], "matern3_2", "constant", False, "BFGS", "LL")
        print(f"group {g} (n = {len(idx)}): alone {km.theta()}, in the batch {th_g[g]}")
        np.testing.assert_allclose(km.theta(), th_g[g], rtol=1e-12)
        assert abs(km.sigma2() - s2_g[g]) <= 1e-12 * s2_g[g]
    # aggregated predict at all 10 000 points
    t0 = time.perf_counter()
    mean, sd = nk.predict(Xn, True)
    t_pred = time.perf_counter() - t0
    print(f"NK predict: {Xn.shape[0]} points in {t_pred:.2f} s = {Xn.shape[0] / t_pred:.0f} points/s, "
          f"fallbacks {nk.predict_fallbacks()}")
    assert np.all(np.isfinite(mean)) and np.all(np.isfinite(sd))
    assert np.all(sd ** 2 <= nk.sigma2() * (1 + 1e-12))
    # interpolation of the design (NestedKrigingTest.cpp:63-75)
    rows = np.random.default_rng(7).choice(X.shape[0], 16, replace=False)
    m_d, s_d = nk.predict(X[rows], True)
    print("interpolation: max |mean - y|", np.max(np.abs(m_d - y[rows])), "max sd", s_d.max())
    np.testing.assert_allclose(m_d, y[rows], atol=1e-3)
    assert s_d.max() < 1e-2


def test_c5_nk_predict_matches_oracle(workload):
    import libkriging_amd as lk
    X, y, group_of, sizes, Xn = workload
    z = np.load(os.path.join(GOLDEN, "c5_nk_predict.npz"))
    assert np.array_equal(z["Xq"], Xn[z["points"]])
    nk = lk.NestedKriging("matern3_2")
    nk.fit_groups(y, X, group_of, "constant", "none", "LL",
                  {"theta": z["theta"], "is_theta_estim": False, "sigma2": float(z["sigma2"]),
                   "is_sigma2_estim": False, "beta": [float(z["beta0"])], "is_beta_estim": False})
    mean, sd = nk.predict(z["Xq"], True)
    e_m = rel_err(mean, z["mean"])
    e_v = np.max(np.abs(sd ** 2 - z["sd"] ** 2)) / float(z["sigma2"])
    print(f"8 points: mean rel err {e_m:.2e}, max |var - var_oracle| / sigma2 {e_v:.2e}, "
          f"fallbacks {nk.predict_fallbacks()} (oracle {int(np.sum(z['info'] > 0))})")
    assert e_m <= 1e-9
    assert e_v <= 1e-9
