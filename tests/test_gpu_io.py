"""Model files (Kriging::save / load, Kriging.cpp:2691-2750; the reference's
JSON version-2 format) and Kriging::update (Kriging.cpp:2433-2633)."""
import base64
import json

import numpy as np
import pytest

from conftest import rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# every field KrigingImpl::load_common_from_json and Kriging::load read
REF_KEYS = ["version", "content", "covType", "X", "centerX", "scaleX", "y", "centerY", "scaleY", "normalize",
            "regmodel", "optim", "objective", "dX", "maxdX", "F", "T", "R", "M", "star", "circ", "z", "Rinv", "beta",
            "est_beta", "theta", "est_theta", "sigma2", "est_sigma2", "noise_model"]


def _arr(node):
    raw = np.frombuffer(base64.b64decode(node["base64_data"]), dtype="<f8")
    if node["type"] == "mat":
        return raw.reshape((node["n_cols"], node["n_rows"])).T
    assert raw.size == node["size"]
    return raw.copy()


def _design(n, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d)) * 2.0 + 1.0
    y = 5.0 + np.sin(3 * X[:, 0]) + X[:, -1] ** 2 + 0.01 * rng.normal(size=n)
    return X, y


def test_save_has_the_reference_fields(tmp_path):
    import libkriging_amd as lk
    X, y = _design(90, 2, 1)
    k = lk.Kriging(y, X, "matern5_2", "linear", True, "BFGS", "LL")
    f = tmp_path / "k.json"
    k.save(f)
    j = json.loads(f.read_text())
    for key in REF_KEYS:
        assert key in j, key
    assert j["version"] == 2 and j["content"] == "Kriging" and j["covType"] == "matern5_2"
    assert j["regmodel"] == "linear" and j["normalize"] is True and j["noise_model"] == "None"
    Xs, ys = _arr(j["X"]), _arr(j["y"])
    assert Xs.shape == (90, 2) and ys.shape == (90,)
    np.testing.assert_allclose(Xs * _arr(j["scaleX"]) + _arr(j["centerX"]), X, rtol=1e-14)
    np.testing.assert_allclose(ys * j["scaleY"] + j["centerY"], y, rtol=1e-14)
    T, R, Rinv = _arr(j["T"]), _arr(j["R"]), _arr(j["Rinv"])
    c = np.linalg.cond(R)
    assert np.allclose(np.triu(T, 1), 0.0)
    assert np.max(np.abs(T @ T.T - R)) < 1e-12
    assert np.max(np.abs(Rinv @ R - np.eye(90))) < 1e-13 * c
    th = _arr(j["theta"])
    np.testing.assert_allclose(th, k.theta().reshape(-1), rtol=0)
    Rk = O.corr_mat("matern5_2", Xs, Xs, th)
    np.fill_diagonal(Rk, 1.0)
    np.testing.assert_allclose(R, Rk, rtol=1e-13, atol=1e-15)
    F, M = _arr(j["F"]), _arr(j["M"])
    assert F.shape == (90, 3) and M.shape == (90, 3)
    np.testing.assert_allclose(M, np.linalg.solve(T, F), rtol=1e-9, atol=1e-11)
    dX = _arr(j["dX"])
    assert dX.shape == (2, 90 * 90)
    np.testing.assert_allclose(dX[:, 3 * 90 + 7], Xs[3] - Xs[7], rtol=0, atol=0)
    np.testing.assert_allclose(dX[:, 7 * 90 + 3], Xs[3] - Xs[7], rtol=0, atol=0)
    np.testing.assert_allclose(_arr(j["maxdX"]), np.max(np.abs(Xs[:, None, :] - Xs[None, :, :]), axis=(0, 1)),
                               rtol=0)
    assert _arr(j["star"]).size == 0 and _arr(j["circ"]).shape == (3, 3)
    assert _arr(j["beta"]).shape == (3,)


@pytest.mark.parametrize("noise", [None, "nugget"])
def test_save_load_roundtrip(tmp_path, noise):
    import libkriging_amd as lk
    X, y = _design(120, 3, 2)
    k = lk.Kriging(y, X, "gauss", "constant", False, "BFGS2", "LL", {}, noise)
    f = tmp_path / "k.json"
    k.save(f)
    k2 = lk.Kriging.load(f)
    Xn = np.random.default_rng(3).uniform(size=(50, 3)) * 2.0 + 1.0
    m1, s1, _, _, _ = k.predict(Xn, True)
    m2, s2, _, _, _ = k2.predict(Xn, True)
    np.testing.assert_allclose(m2, m1, rtol=1e-12, atol=1e-12 * np.max(np.abs(m1)))
    np.testing.assert_allclose(s2, s1, rtol=1e-10, atol=1e-12 * np.max(s1))
    np.testing.assert_array_equal(k2.theta(), k.theta())
    assert k2.sigma2() == k.sigma2() and k2.kernel() == "gauss" and k2.objective() == "LL"
    assert rel_err(k2.logLikelihood(), k.logLikelihood()) < 1e-13
    if noise:
        assert k2.nugget() == k.nugget()


def test_load_a_file_written_elsewhere(tmp_path):
    """a file laid out as another writer would (keys in another order, other
    whitespace, fields this path does not need): load rebuilds the model"""
    import libkriging_amd as lk
    X, y = _design(60, 2, 5)
    ref = lk.Kriging(y, X, "exp", "constant", False, "none", "LL", {"theta": np.array([[0.4, 0.7]])})

    def node_mat(a):
        a = np.asarray(a, dtype="<f8")
        return {"type": "mat", "n_rows": a.shape[0], "n_cols": a.shape[1],
                "base64_data": base64.b64encode(np.asfortranarray(a).T.tobytes()).decode()}

    def node_vec(a, t):
        a = np.asarray(a, dtype="<f8").reshape(-1)
        return {"type": t, "size": a.size, "base64_data": base64.b64encode(a.tobytes()).decode()}

    doc = {"content": "Kriging", "version": 2, "covType": "exp", "X": node_mat(X), "y": node_vec(y, "colvec"),
           "centerX": node_vec([0.0, 0.0], "rowvec"), "scaleX": node_vec([1.0, 1.0], "rowvec"),
           "centerY": 0.0, "scaleY": 1.0, "normalize": False, "regmodel": "constant", "optim": "none",
           "objective": "LL", "theta": node_vec([0.4, 0.7], "colvec"), "est_theta": False,
           "beta": node_vec(ref.beta(), "colvec"), "est_beta": True, "sigma2": ref.sigma2(), "est_sigma2": True,
           "dX": node_mat(np.zeros((2, 0))), "T": node_mat(np.zeros((0, 0))), "nystrom_light": False}
    f = tmp_path / "ext.json"
    f.write_text(json.dumps(doc, indent=2))
    k = lk.Kriging.load(f)
    Xn = np.random.default_rng(6).uniform(size=(30, 2))
    np.testing.assert_allclose(k.predict(Xn, True)[0], ref.predict(Xn, True)[0], rtol=1e-12, atol=1e-13)
    with pytest.raises(RuntimeError, match="requires 2"):
        f2 = tmp_path / "v1.json"
        doc["version"] = 1
        f2.write_text(json.dumps(doc))
        lk.Kriging.load(f2)


def test_update_without_refit_equals_model_on_all_data():
    import libkriging_amd as lk
    X, y = _design(100, 2, 7)
    th = np.array([[0.6, 0.9]])
    k = lk.Kriging(y[:80], X[:80], "matern3_2", "linear", False, "none", "LL", {"theta": th})
    k.update(y[80:], X[80:], False)
    full = lk.Kriging(y, X, "matern3_2", "linear", False, "none", "LL", {"theta": th})
    Xn = np.random.default_rng(8).uniform(size=(40, 2)) * 2.0 + 1.0
    np.testing.assert_allclose(k.predict(Xn, True)[0], full.predict(Xn, True)[0], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(k.predict(Xn, True)[1], full.predict(Xn, True)[1], rtol=1e-9, atol=1e-12)
    assert rel_err(k.sigma2(), full.sigma2()) < 1e-12
    assert k.X().shape == (100, 2)


@pytest.mark.parametrize("noise", [None, "nugget"])
def test_update_without_refit_is_the_block_update_of_the_factor(noise):
    """update(refit=False) with new points keeps the committed factor and
    extends it: T = chol_block(R_new, T_old) (Kriging.cpp:172-187 ->
    LinearAlgebra::update_cholCov, LinearAlgebra.cpp:206-247), checked
    against the oracle's restatement on the same R; M, z and sigma2 follow
    from that factor (numpy solves with the oracle's factor)."""
    import libkriging_amd as lk
    from scipy.linalg import solve_triangular
    X, y = _design(150, 3, 21)
    th = np.array([0.9, 1.1, 1.3])
    par = {"theta": th.reshape(1, -1)}
    alpha = 1.0
    if noise:
        par.update({"sigma2": 0.5, "is_sigma2_estim": False, "nugget": 0.01, "is_nugget_estim": False})
        alpha = 0.5 / 0.51
    k = lk.Kriging(y[:100], X[:100], "matern5_2", "linear", False, "none", "LL", par, noise)
    T_old = k.T()

    def R(Xs):
        Rm = alpha * O.corr_mat("matern5_2", Xs, Xs, th)
        Rm[np.diag_indices(Xs.shape[0])] = 1.0
        return Rm
    To, inc, _ = O.safe_chol_lower(R(X[:100]))
    assert inc == 0
    assert np.max(np.abs(T_old - To)) <= 1e-12 * np.max(np.abs(To))
    k.update(y[100:], X[100:], False)
    L_ref, path = O.chol_block(R(X), To)
    assert path == "block"
    T = k.T()
    err = np.max(np.abs(T - L_ref)) / np.max(np.abs(L_ref))
    print(f"update(noise={noise}): factor rel err vs oracle chol_block {err:.2e}")
    assert err < 1e-10
    F = O.trend_matrix(X, "linear")
    Fs = solve_triangular(L_ref, F, lower=True)
    ys = solve_triangular(L_ref, y, lower=True)
    Q, Rs = np.linalg.qr(Fs)
    beta = solve_triangular(Rs, Q.T @ ys)
    E = ys - Fs @ beta
    assert np.linalg.norm(k.M() - Fs) <= 1e-10 * np.linalg.norm(Fs)
    assert np.linalg.norm(k.z() - E) <= 1e-9 * np.linalg.norm(E)
    np.testing.assert_allclose(k.beta().reshape(-1), beta, rtol=1e-9)
    if not noise:
        assert rel_err(k.sigma2(), E @ E / 150) < 1e-10
    # predictions agree with the model factored from scratch on all the data
    full = lk.Kriging(y, X, "matern5_2", "linear", False, "none", "LL", par, noise)
    Xn = np.random.default_rng(22).uniform(size=(30, 3)) * 2.0 + 1.0
    np.testing.assert_allclose(k.predict(Xn, True)[0], full.predict(Xn, True)[0], rtol=1e-10, atol=1e-12)


def test_update_refit_warm_start_improves_the_likelihood():
    import libkriging_amd as lk
    X, y = _design(130, 2, 9)
    k = lk.Kriging(y[:100], X[:100], "gauss", "constant", True, "BFGS", "LL")
    old = k.theta().reshape(-1).copy()
    k.update(y[100:], X[100:], True)
    new = k.theta().reshape(-1)
    # on the extended data the warm-started optimum is no worse than the old theta
    assert k.logLikelihoodFun(new)[0] >= k.logLikelihoodFun(old)[0] - 1e-9
    assert k.X().shape == (130, 2)


def test_update_refit_nugget_is_a_full_refit():
    import libkriging_amd as lk
    X, y = _design(90, 2, 11)
    k = lk.Kriging(y[:70], X[:70], "matern5_2", "constant", False, "BFGS", "LL", {}, "nugget")
    k.update(y[70:], X[70:], True)
    full = lk.Kriging(y, X, "matern5_2", "constant", False, "BFGS", "LL", {}, "nugget")
    np.testing.assert_allclose(k.theta(), full.theta(), rtol=1e-10)
    assert rel_err(k.nugget(), full.nugget()) < 1e-9


def test_update_rejects_a_wrong_dimension():
    import libkriging_amd as lk
    X, y = _design(40, 2, 13)
    k = lk.Kriging(y, X, "gauss", "constant", False, "none", "LL", {"theta": np.array([[0.5, 0.5]])})
    with pytest.raises(RuntimeError, match="Dimension of new data"):
        k.update(np.ones(3), np.ones((3, 3)), False)


def test_copy_summary_and_getters():
    """lk_kriging_copy (an independent model), summary, covMat
    (KrigingImpl.cpp:60-71, normalised inputs, sigma2 k) and the factor
    getters T, M, z (libkriging_c.h:152-165) against the oracle's
    populate_Model at the committed parameters"""
    import libkriging_amd as lk
    from oracle import oracle as O
    rng = np.random.default_rng(17)
    X = rng.uniform(size=(90, 3)) * np.array([2.0, 1.0, 0.5]) + 1.0
    y = np.sin(2 * X).sum(axis=1)
    k = lk.Kriging(y, X, "matern5_2", "linear", True, "BFGS", "LL")
    c = k.copy()
    assert np.array_equal(c.theta(), k.theta()) and c.sigma2() == k.sigma2()
    assert c.logLikelihood() == k.logLikelihood()
    Xn = rng.uniform(size=(20, 3)) + 1.0
    assert np.array_equal(c.predict(Xn)[0], k.predict(Xn)[0])
    c.update(y[:5] + 0.1, X[:5] + 0.01, False)  # the copy changes, the original does not
    assert k.X().shape[0] == 90 and c.X().shape[0] == 95
    s = k.summary()
    print(s)
    assert "* data (normalized): 90x" in s and "matern5_2" in s and "* trend linear (est.)" in s
    assert "range (est.)" in s and "objective: LL" in s
    # covMat: sigma2 k(x1 - x2) of the normalised inputs
    cm = k.covMat(X[:7], Xn)
    Xs, Xns = (X[:7] - k.centerX()) / k.scaleX(), (Xn - k.centerX()) / k.scaleX()
    ref = k.sigma2() * O.corr_mat("matern5_2", Xs, Xns, k.theta())
    assert np.max(np.abs(cm - ref)) <= 1e-14 * k.sigma2()
    # T, M, z, F and the flags.  The fitted R is ill-conditioned (last
    # pivots ~2e-5): the factors are compared with the oracle's to the spread
    # of valid FP64 orderings (the oracle's loops against LAPACK's dpotrf,
    # times 10, at least 1e-12), and checked by their defining identities
    P = O.OracleProblem(k.X(), k.y(), "matern5_2", "linear")
    m = P.populate_model(k.theta())
    T, M, z = k.T(), k.M(), k.z()
    assert T.shape == (90, 90) and M.shape == (90, 4) and z.shape == (90,)
    R = P.build_R(k.theta())
    F = O.trend_matrix(k.X(), "linear")
    Lnp = np.linalg.cholesky(R)
    spread = np.linalg.norm(Lnp - np.tril(m["L"])) / np.linalg.norm(m["L"])
    bar = max(1e-12, 10.0 * spread)
    print(f"getters: T vs oracle {np.linalg.norm(T - np.tril(m['L'])) / np.linalg.norm(m['L']):.2e}, "
          f"orderings spread {spread:.2e}")
    assert np.linalg.norm(T - np.tril(m["L"])) <= bar * np.linalg.norm(m["L"])
    assert np.max(np.abs(T @ T.T - R)) <= 1e-14 * 90
    assert np.linalg.norm(T @ M - F) <= 1e-13 * np.linalg.norm(F)
    # z = E* = T^-1 (y - F beta): its residual (T has cond ~1e5)
    beta = k.beta().reshape(-1)
    res = T @ z - (k.y() - F @ beta)
    print(f"getters: |T z - (y - F beta)| / |y| = {np.linalg.norm(res) / np.linalg.norm(k.y()):.2e}")
    assert np.linalg.norm(res) <= 1e-10 * np.linalg.norm(k.y())
    assert np.array_equal(k.F(), O.trend_matrix(k.X(), "linear"))
    assert k.is_beta_estim() and k.is_theta_estim() and k.is_sigma2_estim()
    assert k.noise().size == 0
    kf = lk.Kriging(y, X, "gauss", "constant", False, "none", "LL",
                    {"theta": np.array([[0.5, 0.5, 0.5]]), "is_theta_estim": False, "sigma2": 2.0,
                     "is_sigma2_estim": False, "beta": [0.3], "is_beta_estim": False})
    assert not (kf.is_beta_estim() or kf.is_theta_estim() or kf.is_sigma2_estim())
    # beta fixed: z = y* - M beta (Kriging.cpp:2170-2172)
    mf = O.OracleProblem(X, y, "gauss").populate_model(np.array([0.5, 0.5, 0.5]))
    zf = mf["ystar"] - mf["Fstar"][:, 0] * 0.3
    assert np.linalg.norm(kf.z() - zf) <= 1e-10 * np.linalg.norm(zf)
