"""Python surface of the Kriging model, mirroring pylibkriging's
``lk.Kriging`` (= WrappedPyKriging, bindings/Python/pylibkriging/src/
_pylibkriging/pylibkriging.cpp:125-200, Kriging_binding.cpp:57-442): same
constructor forms, keyword names, defaults and return tuples.  Every call goes
through the C ABI of liblkgpu.so (include/lkgpu.h, lk_kriging_*).
"""
import ctypes

import numpy as np

from . import _native as N

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p

N.register("lk_kriging_new", _vp, [_cp, _cp])
N.register("lk_kriging_new_fit", _vp, [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                       _cp, _cp, _cp, ctypes.c_int, _cp, _cp, _dp, ctypes.c_int, _dp,
                                       ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                       ctypes.c_int])
N.register("lk_kriging_delete", None, [_vp])
N.register("lk_kriging_fit", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                            ctypes.c_int, _cp, ctypes.c_int, _cp, _cp])
N.register("lk_kriging_fit_parameters", ctypes.c_int,
           [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _cp, ctypes.c_int, _cp, _cp, _dp,
            ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int])
N.register("lk_kriging_predict", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, _dp, _dp, _dp, _dp, _dp])
N.register("lk_kriging_log_likelihood_fun", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                           _dp, _dp, _dp])
N.register("lk_kriging_log_likelihood_fun_batch", ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int,
                                                                 ctypes.c_int, _dp, _dp])
N.register("lk_kriging_log_likelihood", ctypes.c_double, [_vp])
N.register("lk_kriging_leave_one_out_fun", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp])
N.register("lk_kriging_leave_one_out", ctypes.c_double, [_vp])
N.register("lk_kriging_log_marg_post_fun", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp])
N.register("lk_kriging_log_marg_post", ctypes.c_double, [_vp])
N.register("lk_kriging_update", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int])
N.register("lk_kriging_save", ctypes.c_int, [_vp, _cp])
N.register("lk_kriging_load", _vp, [_cp])
N.register("lk_kriging_leave_one_out_vec", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, _dp])
for _name in ("lk_kriging_kernel", "lk_kriging_optim", "lk_kriging_objective", "lk_kriging_regmodel",
              "lk_kriging_noise_model"):
    N.register(_name, _cp, [_vp])
N.register("lk_kriging_is_normalize", ctypes.c_int, [_vp])
N.register("lk_kriging_get_X", ctypes.c_int, [_vp, _dp, _ip, _ip])
for _name in ("lk_kriging_get_centerX", "lk_kriging_get_scaleX", "lk_kriging_get_y", "lk_kriging_get_beta",
              "lk_kriging_get_theta"):
    N.register(_name, ctypes.c_int, [_vp, _dp, _ip])
for _name in ("lk_kriging_get_centerY", "lk_kriging_get_scaleY", "lk_kriging_get_sigma2", "lk_kriging_get_nugget"):
    N.register(_name, ctypes.c_double, [_vp])
N.register("lk_kriging_fit_statistics", ctypes.c_int, [_vp, _ip, _ip])
N.register("lk_kriging_set_batched_fit", ctypes.c_int, [_vp, ctypes.c_int])
N.register("lk_kriging_copy", _vp, [_vp])
N.register("lk_kriging_summary", ctypes.c_char_p, [_vp])
N.register("lk_kriging_cov_mat", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                                _dp])
for _name in ("lk_kriging_get_F", "lk_kriging_get_T", "lk_kriging_get_M"):
    N.register(_name, ctypes.c_int, [_vp, _dp, _ip, _ip])
for _name in ("lk_kriging_get_z", "lk_kriging_get_noise"):
    N.register(_name, ctypes.c_int, [_vp, _dp, _ip])
for _name in ("lk_kriging_is_beta_estim", "lk_kriging_is_theta_estim", "lk_kriging_is_sigma2_estim",
              "lk_kriging_is_nugget_estim"):
    N.register(_name, ctypes.c_int, [_vp])
N.register("lk_optim_set", ctypes.c_int, [_cp, ctypes.c_double])
N.register("lk_optim_get", ctypes.c_int, [_cp, _dp])
N.register("lk_linalg_set", ctypes.c_int, [_cp, ctypes.c_double])
N.register("lk_linalg_get", ctypes.c_int, [_cp, _dp])
N.register("lk_linalg_chol_safe", ctypes.c_int, [_dp, ctypes.c_int, _dp, _ip])
for _name in ("lk_linalg_rcond_chol", "lk_linalg_rcond_approx_chol"):
    N.register(_name, ctypes.c_int, [_dp, ctypes.c_int, _dp])
N.register("lk_linalg_chol_block", ctypes.c_int, [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp])
N.register("lk_random_reset_seed", ctypes.c_int, [ctypes.c_uint])
LK_OBJECTIVE_FN = ctypes.CFUNCTYPE(ctypes.c_double, _dp, _dp, _vp)
N.register("lk_lbfgsb_minimize", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, LK_OBJECTIVE_FN, _vp, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _ip, _ip,
                                                ctypes.c_char_p, ctypes.c_int])


def _b(s):
    return None if s is None else str(s).encode()


def _col(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _mat(X):
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X.reshape(-1, 1)
    return np.asfortranarray(X)


def _params(parameters, d):
    """parameters dict keys of Kriging_binding.cpp:39-48 -> C arguments."""
    p = dict(parameters or {})
    sig = p.get("sigma2")
    th = p.get("theta")
    be = p.get("beta")
    nu = p.get("nugget")
    sig_c = ctypes.byref(ctypes.c_double(float(sig))) if sig is not None else None
    nu_c = ctypes.byref(ctypes.c_double(float(nu))) if nu is not None else None
    th_a = None
    if th is not None:
        th = np.asarray(th, dtype=np.float64)
        if th.ndim == 1:
            th = th.reshape(1, -1)
        if th.shape[1] != d and th.shape[0] == d:  # fit_setup_impl:833-834 transposes a d x k matrix
            th = th.T
        th_a = np.ascontiguousarray(th)  # rows x d, row-major
    be_a = _col(be) if be is not None else None
    return dict(sig=sig_c, sig_e=int(p.get("is_sigma2_estim", True)),
                th=th_a, th_e=int(p.get("is_theta_estim", True)),
                be=be_a, be_e=int(p.get("is_beta_estim", True)),
                nu=nu_c, nu_e=int(p.get("is_nugget_estim", True)))


class Kriging:
    """Kriging(kernel) | Kriging(kernel, noise_model) |
    Kriging(y, X, kernel, regmodel="constant", normalize=False, optim="BFGS",
            objective="LL", parameters={}, noise=None)"""

    def __init__(self, *args, **kwargs):
        self._h = None
        self._keep = None
        if len(args) + len(kwargs) <= 2 and "y" not in kwargs and (not args or isinstance(args[0], str)):
            kernel = kwargs.get("kernel", args[0] if args else None)
            noise = kwargs.get("noise_model", args[1] if len(args) > 1 else None)
            h = N.lib().lk_kriging_new(_b(kernel), _b(noise if noise is not None else "none"))
            if not h:
                raise ValueError(N.lib().lk_get_last_error().decode())
            self._h = h
            return
        names = ["y", "X", "kernel", "regmodel", "normalize", "optim", "objective", "parameters", "noise"]
        vals = dict(zip(names, args))
        vals.update(kwargs)
        y, X, kernel = vals["y"], vals["X"], vals["kernel"]
        noise = vals.get("noise")
        if noise is not None and not isinstance(noise, str):
            raise ValueError("heterogeneous noise is not supported on this path")
        if isinstance(noise, str) and noise != "nugget":
            raise ValueError(f"noise string must be 'nugget', got '{noise}'")
        h = N.lib().lk_kriging_new(_b(kernel), _b(noise or "none"))
        if not h:
            raise ValueError(N.lib().lk_get_last_error().decode())
        self._h = h
        self.fit(y, X, vals.get("regmodel", "constant"), vals.get("normalize", False), vals.get("optim", "BFGS"),
                 vals.get("objective", "LL"), vals.get("parameters", None))

    @classmethod
    def _borrowed(cls, h, owner):
        """a view of a model owned by another object (kept alive through owner)"""
        k = cls.__new__(cls)
        k._h = h
        k._owned = False
        k._keep = owner
        return k

    def __del__(self):
        try:
            if self._h and getattr(self, "_owned", True):
                N.lib().lk_kriging_delete(self._h)
                self._h = None
        except Exception:
            pass

    def fit(self, y, X, regmodel="constant", normalize=False, optim="BFGS", objective="LL", parameters=None,
            noise=None):
        if noise is not None and noise != "nugget":
            raise ValueError("heterogeneous noise is not supported on this path")
        y = _col(y)
        X = _mat(X)
        n, d = X.shape
        P = _params(parameters, d)
        th = P["th"]
        self._keep = (y, X, th, P["be"])
        N.check(N.lib().lk_kriging_fit_parameters(
            self._h, N.ptr(y), y.size, N.ptr(X), n, d, _b(regmodel), int(bool(normalize)), _b(optim),
            _b(objective), P["sig"], P["sig_e"], N.ptr(th) if th is not None else None,
            th.size if th is not None else 0, P["th_e"], N.ptr(P["be"]) if P["be"] is not None else None,
            P["be"].size if P["be"] is not None else 0, P["be_e"], P["nu"], P["nu_e"]))

    def predict(self, X, return_stdev=True, return_cov=False, return_deriv=False):
        """(mean, stdev, cov, mean_deriv, stdev_deriv) as Kriging::predict
        (Kriging.cpp:2241-2285); unrequested outputs are empty"""
        X = _mat(X)
        m, d = X.shape
        mean = np.zeros(m)
        sd = np.zeros(m) if return_stdev else None
        cov = np.zeros((m, m), order="F") if return_cov else None
        dm = np.zeros((m, d), order="F") if return_deriv else None
        ds = np.zeros((m, d), order="F") if return_deriv else None

        def p_(a):
            return N.ptr(a) if a is not None else None
        N.check(N.lib().lk_kriging_predict(self._h, N.ptr(X), m, d, int(return_stdev), int(return_cov),
                                           int(return_deriv), N.ptr(mean), p_(sd), p_(cov), p_(dm), p_(ds)))
        empty = np.zeros((0, 0))
        return (mean, (sd if sd is not None else np.zeros(0)), cov if cov is not None else empty,
                dm if dm is not None else empty, ds if ds is not None else empty)

    def logLikelihoodFun(self, theta, return_grad=False, want_hess=False):
        th = _col(theta)
        ll = ctypes.c_double()
        grad = np.zeros(th.size) if return_grad else None
        N.check(N.lib().lk_kriging_log_likelihood_fun(self._h, N.ptr(th), th.size, int(bool(return_grad)), 0,
                                                      ctypes.byref(ll), N.ptr(grad) if grad is not None else None,
                                                      None))
        return ll.value, (grad if grad is not None else np.zeros(0)), np.zeros((0, 0))

    def logLikelihoodFunBatch(self, thetas, return_grad=True):
        """Several parameter vectors (rows) in one batched GPU pipeline."""
        th = np.ascontiguousarray(np.atleast_2d(np.asarray(thetas, dtype=np.float64)))
        nb, gn = th.shape
        ll = np.zeros(nb)
        grad = np.zeros((nb, gn)) if return_grad else None
        N.check(N.lib().lk_kriging_log_likelihood_fun_batch(self._h, nb, th.ctypes.data_as(_dp), gn,
                                                            int(bool(return_grad)), N.ptr(ll),
                                                            grad.ctypes.data_as(_dp) if grad is not None else None))
        return ll, grad

    def update(self, y_u, X_u, refit=True, noise_u=None):
        """Kriging::update (Kriging.cpp:2433-2633): append the points (X_u, y_u)"""
        if noise_u is not None:
            raise ValueError("heterogeneous noise is not supported on this path")
        y = _col(y_u)
        X = _mat(X_u)
        n, d = X.shape
        N.check(N.lib().lk_kriging_update(self._h, N.ptr(y), y.size, None, 0, N.ptr(X), n, d, int(bool(refit))))

    def save(self, filename):
        """the reference's JSON version-2 model file (Kriging::save)"""
        N.check(N.lib().lk_kriging_save(self._h, _b(str(filename))))

    @classmethod
    def load(cls, filename):
        """a model from a JSON version-2 file (Kriging::load)"""
        h = N.lib().lk_kriging_load(_b(str(filename)))
        if not h:
            raise RuntimeError(N.lib().lk_get_last_error().decode())
        k = cls.__new__(cls)
        k._h = h
        k._owned = True
        k._keep = None
        return k

    def leaveOneOutFun(self, theta, return_grad=False):
        """(loo, grad) as Kriging::leaveOneOutFun (Kriging.cpp:470-476)"""
        th = _col(theta)
        loo = ctypes.c_double()
        grad = np.zeros(th.size) if return_grad else None
        N.check(N.lib().lk_kriging_leave_one_out_fun(self._h, N.ptr(th), th.size, int(bool(return_grad)),
                                                     ctypes.byref(loo), N.ptr(grad) if grad is not None else None))
        return loo.value, (grad if grad is not None else np.zeros(0))

    def leaveOneOutVec(self, theta):
        """(yhat, stderr) of the n leave-one-out predictions (Kriging.cpp:478-485)"""
        th = _col(theta)
        nn = ctypes.c_int()
        N.lib().lk_kriging_get_X(self._h, None, ctypes.byref(nn), None)
        n = nn.value
        yhat = np.zeros(n)
        se = np.zeros(n)
        N.check(N.lib().lk_kriging_leave_one_out_vec(self._h, N.ptr(th), th.size, N.ptr(yhat), N.ptr(se)))
        return yhat, se

    def logMargPostFun(self, theta, return_grad=False):
        """(lmp, grad) as Kriging::logMargPostFun (Kriging.cpp:652-660)"""
        th = _col(theta)
        v = ctypes.c_double()
        grad = np.zeros(th.size) if return_grad else None
        N.check(N.lib().lk_kriging_log_marg_post_fun(self._h, N.ptr(th), th.size, int(bool(return_grad)),
                                                     ctypes.byref(v), N.ptr(grad) if grad is not None else None))
        return v.value, (grad if grad is not None else np.zeros(0))

    def logMargPost(self):
        v = N.lib().lk_kriging_log_marg_post(self._h)
        if np.isnan(v):
            N.check(-1)
        return v

    def leaveOneOut(self):
        v = N.lib().lk_kriging_leave_one_out(self._h)
        if np.isnan(v):
            N.check(-1)
        return v

    def logLikelihood(self):
        v = N.lib().lk_kriging_log_likelihood(self._h)
        if np.isnan(v):
            err = N.lib().lk_get_last_error().decode()
            if err:
                raise RuntimeError(err)
        return v

    def _vec(self, fn):
        n = ctypes.c_int()
        getattr(N.lib(), fn)(self._h, None, ctypes.byref(n))
        out = np.zeros(n.value)
        getattr(N.lib(), fn)(self._h, N.ptr(out), ctypes.byref(n))
        return out

    def theta(self):
        return self._vec("lk_kriging_get_theta")

    def beta(self):
        return self._vec("lk_kriging_get_beta")

    def sigma2(self):
        return N.lib().lk_kriging_get_sigma2(self._h)

    def nugget(self):
        return N.lib().lk_kriging_get_nugget(self._h)

    def centerX(self):
        return self._vec("lk_kriging_get_centerX")

    def scaleX(self):
        return self._vec("lk_kriging_get_scaleX")

    def centerY(self):
        return N.lib().lk_kriging_get_centerY(self._h)

    def scaleY(self):
        return N.lib().lk_kriging_get_scaleY(self._h)

    def y(self):
        return self._vec("lk_kriging_get_y")

    def X(self):
        n, d = ctypes.c_int(), ctypes.c_int()
        N.lib().lk_kriging_get_X(self._h, None, ctypes.byref(n), ctypes.byref(d))
        out = np.zeros((n.value, d.value), order="F")
        N.lib().lk_kriging_get_X(self._h, N.ptr(out), ctypes.byref(n), ctypes.byref(d))
        return out

    def kernel(self):
        return N.lib().lk_kriging_kernel(self._h).decode()

    def optim(self):
        return N.lib().lk_kriging_optim(self._h).decode()

    def objective(self):
        return N.lib().lk_kriging_objective(self._h).decode()

    def regmodel(self):
        return N.lib().lk_kriging_regmodel(self._h).decode()

    def noise_model(self):
        return N.lib().lk_kriging_noise_model(self._h).decode()

    def normalize(self):
        return bool(N.lib().lk_kriging_is_normalize(self._h))

    def fit_statistics(self):
        e, b = ctypes.c_int(), ctypes.c_int()
        N.check(N.lib().lk_kriging_fit_statistics(self._h, ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def copy(self):
        """an independent model with this one's state (pylibkriging copy)"""
        h = N.lib().lk_kriging_copy(self._h)
        if not h:
            raise RuntimeError(N.lib().lk_get_last_error().decode())
        k = Kriging.__new__(Kriging)
        k._h = h
        k._keep = None
        return k

    def summary(self):
        s = N.lib().lk_kriging_summary(self._h)
        if s is None:
            raise RuntimeError(N.lib().lk_get_last_error().decode())
        return s.decode()

    def __repr__(self):
        return self.summary()

    def covMat(self, X1, X2):
        """sigma2 k(x1_i - x2_j) (KrigingImpl::covMat), n1 x n2"""
        X1, X2 = _mat(X1), _mat(X2)
        out = np.zeros((X1.shape[0], X2.shape[0]), order="F")
        N.check(N.lib().lk_kriging_cov_mat(self._h, N.ptr(X1), X1.shape[0], X1.shape[1], N.ptr(X2), X2.shape[0],
                                           X2.shape[1], N.ptr(out)))
        return out

    def _mat_get(self, fn):
        n, d = ctypes.c_int(), ctypes.c_int()
        N.check(getattr(N.lib(), fn)(self._h, None, ctypes.byref(n), ctypes.byref(d)))
        out = np.zeros((n.value, d.value), order="F")
        N.check(getattr(N.lib(), fn)(self._h, N.ptr(out), ctypes.byref(n), ctypes.byref(d)))
        return out

    def F(self):
        return self._mat_get("lk_kriging_get_F")

    def T(self):
        return self._mat_get("lk_kriging_get_T")

    def M(self):
        return self._mat_get("lk_kriging_get_M")

    def z(self):
        n = ctypes.c_int()
        N.check(N.lib().lk_kriging_get_z(self._h, None, ctypes.byref(n)))
        out = np.zeros(n.value)
        N.check(N.lib().lk_kriging_get_z(self._h, N.ptr(out), ctypes.byref(n)))
        return out

    def noise(self):
        n = ctypes.c_int()
        N.check(N.lib().lk_kriging_get_noise(self._h, None, ctypes.byref(n)))
        return np.zeros(n.value)

    def is_beta_estim(self):
        return bool(N.lib().lk_kriging_is_beta_estim(self._h))

    def is_theta_estim(self):
        return bool(N.lib().lk_kriging_is_theta_estim(self._h))

    def is_sigma2_estim(self):
        return bool(N.lib().lk_kriging_is_sigma2_estim(self._h))

    def is_nugget_estim(self):
        return bool(N.lib().lk_kriging_is_nugget_estim(self._h))

    def set_batched_fit(self, batched):
        N.check(N.lib().lk_kriging_set_batched_fit(self._h, int(bool(batched))))


class Optim:
    """Static optimisation settings (pylibkriging.cpp:448-469)."""

    @staticmethod
    def _set(name, v):
        N.check(N.lib().lk_optim_set(_b(name), float(v)))

    @staticmethod
    def _get(name):
        v = ctypes.c_double()
        N.check(N.lib().lk_optim_get(_b(name), ctypes.byref(v)))
        return v.value

    is_reparametrized = staticmethod(lambda: bool(Optim._get("reparametrize")))
    use_reparametrize = staticmethod(lambda b: Optim._set("reparametrize", bool(b)))
    get_theta_lower_factor = staticmethod(lambda: Optim._get("theta_lower_factor"))
    set_theta_lower_factor = staticmethod(lambda v: Optim._set("theta_lower_factor", v))
    get_theta_upper_factor = staticmethod(lambda: Optim._get("theta_upper_factor"))
    set_theta_upper_factor = staticmethod(lambda v: Optim._set("theta_upper_factor", v))
    variogram_bounds_heuristic_used = staticmethod(lambda: bool(Optim._get("variogram_bounds_heuristic")))
    use_variogram_bounds_heuristic = staticmethod(lambda b: Optim._set("variogram_bounds_heuristic", bool(b)))
    get_log_level = staticmethod(lambda: int(Optim._get("log_level")))
    set_log_level = staticmethod(lambda v: Optim._set("log_level", v))
    get_max_iteration = staticmethod(lambda: int(Optim._get("max_iteration")))
    set_max_iteration = staticmethod(lambda v: Optim._set("max_iteration", v))
    get_gradient_tolerance = staticmethod(lambda: Optim._get("gradient_tolerance"))
    set_gradient_tolerance = staticmethod(lambda v: Optim._set("gradient_tolerance", v))
    get_objective_rel_tolerance = staticmethod(lambda: Optim._get("objective_rel_tolerance"))
    set_objective_rel_tolerance = staticmethod(lambda v: Optim._set("objective_rel_tolerance", v))


def lbfgsb_minimize(fun, x0, lower, upper, max_iter=20, pgtol=1e-3, factr=1e10, m=5):
    """Run the fit's L-BFGS-B on fun(x) -> (f, g).  Returns (x, f, iters,
    evals, task)."""
    x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).copy())
    lo = np.ascontiguousarray(np.asarray(lower, dtype=np.float64))
    up = np.ascontiguousarray(np.asarray(upper, dtype=np.float64))
    n = x.size
    err = []

    def cb(xp, gp, _user):
        try:
            xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
            f, g = fun(xv)
            ga = np.ctypeslib.as_array(gp, shape=(n,))
            ga[:] = np.asarray(g, dtype=np.float64)
            return float(f)
        except Exception as e:  # keep the C loop alive, re-raise afterwards
            err.append(e)
            return float("nan")

    cfun = LK_OBJECTIVE_FN(cb)
    f = ctypes.c_double()
    it = ctypes.c_int()
    ev = ctypes.c_int()
    task = ctypes.create_string_buffer(128)
    N.check(N.lib().lk_lbfgsb_minimize(n, x.ctypes.data_as(_dp), lo.ctypes.data_as(_dp), up.ctypes.data_as(_dp),
                                       cfun, None, int(max_iter), float(pgtol), float(factr), int(m),
                                       ctypes.byref(f), ctypes.byref(it), ctypes.byref(ev), task, 128))
    if err:
        raise err[0]
    return x, f.value, it.value, ev.value, task.value.decode()


def set_num_nugget(v):
    N.check(N.lib().lk_linalg_set(b"num_nugget", float(v)))


def check_chol_rcond(b):
    N.check(N.lib().lk_linalg_set(b"chol_rcond_check", float(bool(b))))


def _linalg_get(name):
    v = ctypes.c_double()
    N.check(N.lib().lk_linalg_get(name, ctypes.byref(v)))
    return v.value


def get_num_nugget():
    """LinearAlgebra::get_num_nugget (LinearAlgebra.cpp:37-39)."""
    return _linalg_get(b"num_nugget")


def chol_rcond_checked():
    """LinearAlgebra::chol_rcond_checked (LinearAlgebra.cpp:59-61)."""
    return bool(_linalg_get(b"chol_rcond_check"))


def _square(M, name):
    M = np.asfortranarray(np.asarray(M, dtype=np.float64))
    if M.ndim != 2 or M.shape[0] != M.shape[1] or M.shape[0] == 0:
        raise ValueError(f"{name} must be a non-empty square matrix")
    return M


def chol_safe(X, return_inc=False):
    """LinearAlgebra::safe_chol_lower (LinearAlgebra.cpp:43-98) on the device:
    the lower factor of X plus the retry ladder's jitter (R binding
    linalg_chol_safe, bindings/R/rlibkriging/src/linalg.cpp:39-43)."""
    X = _square(X, "X")
    n = X.shape[0]
    L = np.zeros((n, n), order="F")
    inc = ctypes.c_int()
    N.check(N.lib().lk_linalg_chol_safe(X.ctypes.data_as(_dp), n, L.ctypes.data_as(_dp), ctypes.byref(inc)))
    return (L, inc.value) if return_inc else L


def rcond_chol(L):
    """LinearAlgebra::rcond_chol (LinearAlgebra.cpp:106-115): rcond(L)^2 of a
    lower factor, 1-norm estimate (R binding linalg_rcond_chol)."""
    L = _square(L, "L")
    r = ctypes.c_double()
    N.check(N.lib().lk_linalg_rcond_chol(L.ctypes.data_as(_dp), L.shape[0], ctypes.byref(r)))
    return r.value


def rcond_approx_chol(L):
    """LinearAlgebra::rcond_approx_chol (LinearAlgebra.cpp:120-139)."""
    L = _square(L, "L")
    r = ctypes.c_double()
    N.check(N.lib().lk_linalg_rcond_approx_chol(L.ctypes.data_as(_dp), L.shape[0], ctypes.byref(r)))
    return r.value


def chol_block(C, Loo):
    """LinearAlgebra::chol_block (LinearAlgebra.cpp:257-299): the lower factor
    of C given the factor Loo of its leading block (R binding
    linalg_chol_block)."""
    C = _square(C, "C")
    Loo = _square(Loo, "Loo")
    n, no = C.shape[0], Loo.shape[0]
    if no > n:
        raise ValueError("Loo must not be larger than C")
    L = np.zeros((n, n), order="F")
    N.check(N.lib().lk_linalg_chol_block(C.ctypes.data_as(_dp), n, Loo.ctypes.data_as(_dp), no,
                                         L.ctypes.data_as(_dp)))
    return L


def reset_seed(seed):
    N.check(N.lib().lk_random_reset_seed(int(seed)))
