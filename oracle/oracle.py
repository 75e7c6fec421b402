"""ctypes front-end of the CPU oracle (oracle/lk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product package
(libkriging_amd) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblk_oracle.so")

KERNELS = {"gauss": 0, "exp": 1, "matern3_2": 2, "matern5_2": 3}
NOISE_NONE, NOISE_NUGGET = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class CholCfg(ctypes.Structure):
    _fields_ = [("num_nugget", ctypes.c_double), ("rcond_check", ctypes.c_int),
                ("max_inc", ctypes.c_int), ("min_rcond", ctypes.c_double)]


class Problem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("d", ctypes.c_int), ("p", ctypes.c_int),
                ("X", _dp), ("y", _dp), ("F", _dp),
                ("kernel", ctypes.c_int), ("noise", ctypes.c_int),
                ("est_sigma2", ctypes.c_int), ("est_nugget", ctypes.c_int),
                ("est_beta", ctypes.c_int),
                ("sigma2", ctypes.c_double), ("nugget", ctypes.c_double),
                ("alpha", ctypes.c_double)]


class Model(ctypes.Structure):
    _fields_ = [("R", _dp), ("L", _dp), ("Rinv", _dp), ("Fstar", _dp), ("ystar", _dp), ("Rstar", _dp),
                ("Estar", _dp), ("betahat", _dp), ("SSEstar", ctypes.c_double), ("inc_cond", ctypes.c_int),
                ("rcond", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: run `make -C {_HERE}`")
        L = ctypes.CDLL(_LIB_PATH)
        L.lko_cov.restype = ctypes.c_double
        L.lko_cov.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.lko_dlncov_dtheta.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.lko_corr_mat.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.lko_corr_mat.restype = None
        L.lko_build_R.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                  ctypes.c_double, _dp, _dp]
        L.lko_potrf_lower.restype = ctypes.c_int
        L.lko_potrf_lower.argtypes = [_dp, ctypes.c_int]
        L.lko_trcon_lower.restype = ctypes.c_double
        L.lko_trcon_lower.argtypes = [_dp, ctypes.c_int]
        L.lko_safe_chol_lower.restype = ctypes.c_int
        L.lko_safe_chol_lower.argtypes = [_dp, ctypes.c_int, ctypes.POINTER(CholCfg), _dp, _ip, _dp]
        L.lko_inv_sympd.argtypes = [_dp, ctypes.c_int, _dp]
        L.lko_solve_lower.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.lko_solve_lower.restype = None
        L.lko_log_likelihood.restype = ctypes.c_int
        L.lko_log_likelihood.argtypes = [ctypes.POINTER(Problem), _dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(CholCfg), _dp, _dp, _ip]
        L.lko_theta_bounds.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.lko_predict.restype = ctypes.c_int
        L.lko_predict.argtypes = [ctypes.POINTER(Problem), _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(CholCfg), _dp, ctypes.c_int, _dp, _dp]
        L.lko_populate_model.restype = ctypes.c_int
        L.lko_populate_model.argtypes = [ctypes.POINTER(Problem), _dp, ctypes.c_double, ctypes.POINTER(CholCfg),
                                         ctypes.POINTER(Model)]
        L.lko_solve_upper_t.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.lko_solve_upper_t.restype = None
        L.lko_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def use_lapack(enable=True):
    """Route the oracle's dpotrf / dtrtrs / dtrcon through scipy's LAPACK
    (the routines the reference reaches through Armadillo) or back to the
    restated loops.  BLAS threads are those of scipy's OpenBLAS
    (threadpoolctl limits them)."""
    L = lib()
    L.lko_set_lapack.argtypes = [ctypes.c_void_p] * 3
    L.lko_set_lapack.restype = None
    if not enable:
        L.lko_set_lapack(None, None, None)
        return
    import scipy.linalg.cython_lapack as cl
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    gname = ctypes.pythonapi.PyCapsule_GetName
    gname.restype = ctypes.c_char_p
    gname.argtypes = [ctypes.py_object]
    ptrs = []
    for name in ("dpotrf", "dtrtrs", "dtrcon"):
        cap = cl.__pyx_capi__[name]
        ptrs.append(get(cap, gname(cap)))
    L.lko_set_lapack(*ptrs)


def default_cfg(num_nugget=1e-10, rcond_check=True, max_inc=10, min_rcond=1e-18):
    return CholCfg(num_nugget, int(rcond_check), max_inc, min_rcond)


def trend_matrix(X, regmodel):
    """Trend::regressionModelMatrix (Trend.cpp:34-61) for constant / linear."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    if regmodel == "constant":
        return np.ones((n, 1))
    if regmodel == "linear":
        return np.hstack([np.ones((n, 1)), X])
    if regmodel == "none":
        return np.ones((n, 0))
    raise ValueError(regmodel)


class OracleProblem:
    """Keeps the arrays alive while the C struct points at them."""

    def __init__(self, X, y, kernel="gauss", regmodel="constant", noise=None,
                 est_sigma2=True, sigma2=1.0, est_nugget=True, nugget=0.0, alpha=1.0,
                 est_beta=True):
        X = np.asarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X.reshape(-1, 1)
        self.X = _f64(X)
        self.y = _f64(np.asarray(y, dtype=np.float64).reshape(-1))
        self.F = _f64(trend_matrix(self.X, regmodel))
        n, d = self.X.shape
        self.n, self.d, self.p = n, d, self.F.shape[1]
        self.kernel = kernel
        self.noise = noise
        self.s = Problem(n, d, self.p, _ptr(self.X), _ptr(self.y), _ptr(self.F),
                         KERNELS[kernel], NOISE_NUGGET if noise == "nugget" else NOISE_NONE,
                         int(est_sigma2), int(est_nugget), int(est_beta),
                         float(sigma2), float(nugget), float(alpha))

    def log_likelihood(self, gamma, want_grad=True, cfg=None):
        gamma = _f64(np.asarray(gamma, dtype=np.float64).reshape(-1))
        cfg = cfg or default_cfg()
        ll = ctypes.c_double()
        grad = np.zeros(gamma.size)
        inc = ctypes.c_int(0)
        rc = lib().lko_log_likelihood(ctypes.byref(self.s), _ptr(gamma), gamma.size, int(want_grad),
                                      ctypes.byref(cfg), ctypes.byref(ll), _ptr(grad), ctypes.byref(inc))
        if rc != 0:
            raise RuntimeError(lib().lko_last_error().decode())
        return ll.value, grad, inc.value

    def populate_model(self, theta, alpha=1.0, cfg=None):
        """KrigingImpl::populate_Model (KrigingImpl.cpp:73-125): dict of R
        (un-jittered), L, Rinv, Fstar, ystar, Rstar, Estar, betahat, SSEstar,
        inc, rcond, and x = L^-T E* (Kriging.cpp:294)"""
        theta = _f64(np.asarray(theta, dtype=np.float64).reshape(-1))
        n, p = self.n, self.p
        arr = {k: np.zeros((n, n), order="F") for k in ("R", "L", "Rinv")}
        arr["Fstar"] = np.zeros((n, max(p, 1)), order="F")
        arr["ystar"] = np.zeros(n)
        arr["Rstar"] = np.zeros((max(p, 1), max(p, 1)), order="F")
        arr["Estar"] = np.zeros(n)
        arr["betahat"] = np.zeros(max(p, 1))
        m = Model(*[_ptr(arr[k]) for k in ("R", "L", "Rinv", "Fstar", "ystar", "Rstar", "Estar", "betahat")],
                  0.0, 0, 0.0)
        rc = lib().lko_populate_model(ctypes.byref(self.s), _ptr(theta), float(alpha), ctypes.byref(cfg or default_cfg()),
                                      ctypes.byref(m))
        if rc != 0:
            raise RuntimeError(lib().lko_last_error().decode())
        x = arr["Estar"].copy()
        lib().lko_solve_upper_t(_ptr(arr["L"]), n, _ptr(x), 1)
        out = dict(arr)
        out["Fstar"] = arr["Fstar"][:, :p]
        out["Rstar"] = arr["Rstar"][:p, :p]
        out["betahat"] = arr["betahat"][:p]
        out.update(SSEstar=m.SSEstar, inc=m.inc_cond, rcond=m.rcond, x=x)
        return out

    def build_R(self, theta, factor=1.0):
        theta = _f64(theta)
        R = np.zeros((self.n, self.n), order="F")
        lib().lko_build_R(KERNELS[self.kernel], _ptr(self.X), self.n, self.d, _ptr(theta),
                          float(factor), None, _ptr(R))
        return R

    def predict(self, theta, alpha, sigma2, Xn, cfg=None):
        theta = _f64(theta)
        Xn = _f64(np.asarray(Xn, dtype=np.float64).reshape(-1, self.d))
        m = Xn.shape[0]
        mean = np.zeros(m)
        sd = np.zeros(m)
        cfg = cfg or default_cfg()
        rc = lib().lko_predict(ctypes.byref(self.s), _ptr(theta), float(alpha), float(sigma2),
                               ctypes.byref(cfg), _ptr(Xn), m, _ptr(mean), _ptr(sd))
        if rc != 0:
            raise RuntimeError(lib().lko_last_error().decode())
        return mean, sd


def cov(kernel, dx, theta):
    dx = _f64(dx)
    theta = _f64(theta)
    return lib().lko_cov(KERNELS[kernel], _ptr(dx), _ptr(theta), dx.size)


def corr_mat(kernel, X1, X2, theta):
    """NestedKriging::corrMat: k(x1_i - x2_j), n1 x n2"""
    X1 = _f64(np.asarray(X1, dtype=np.float64))
    X2 = _f64(np.asarray(X2, dtype=np.float64))
    theta = _f64(theta)
    out = np.zeros((X1.shape[0], X2.shape[0]), order="F")
    lib().lko_corr_mat(KERNELS[kernel], _ptr(X1), X1.shape[0], _ptr(X2), X2.shape[0], X1.shape[1], _ptr(theta),
                       _ptr(out))
    return out


def dlncov_dtheta(kernel, dx, theta):
    dx = _f64(dx)
    theta = _f64(theta)
    out = np.zeros(dx.size)
    lib().lko_dlncov_dtheta(KERNELS[kernel], _ptr(dx), _ptr(theta), dx.size, _ptr(out))
    return out


def potrf_lower(A):
    L = _f64(np.array(A, dtype=np.float64))
    info = lib().lko_potrf_lower(_ptr(L), L.shape[0])
    return L, info


def trcon_lower(L):
    L = _f64(L)
    return lib().lko_trcon_lower(_ptr(L), L.shape[0])


def safe_chol_lower(R, cfg=None):
    R = _f64(R)
    n = R.shape[0]
    L = np.zeros((n, n), order="F")
    inc = ctypes.c_int(0)
    rc_ = ctypes.c_double(0)
    cfg = cfg or default_cfg()
    rc = lib().lko_safe_chol_lower(_ptr(R), n, ctypes.byref(cfg), _ptr(L), ctypes.byref(inc),
                                   ctypes.byref(rc_))
    if rc != 0:
        raise RuntimeError(lib().lko_last_error().decode())
    return L, inc.value, rc_.value


def solve_lower(L, B):
    """Loo^-1 B by forward substitution (LinearAlgebra::solve on a lower
    factor, LinearAlgebra.cpp:385-387; LAPACK dtrtrs under use_lapack())."""
    L = _f64(L)
    B = _f64(np.array(B, dtype=np.float64).reshape(L.shape[0], -1))
    lib().lko_solve_lower(_ptr(L), L.shape[0], _ptr(B), B.shape[1])
    return B


def chol_block(C, Loo, cfg=None):
    """LinearAlgebra::chol_block (LinearAlgebra.cpp:257-299): the lower factor
    of C from the factor Loo of its leading no x no block.  L_ou' = Loo^-1
    C_ou, L_uu = safe_chol(C_uu - L_ou L_ou'); when that ladder fails, the
    factor of C from scratch.  Returns (L, path) with path "block" or
    "full"."""
    C = _f64(C)
    Loo = np.tril(_f64(Loo))
    n, no = C.shape[0], Loo.shape[0]
    L = np.zeros((n, n), order="F")
    L[:no, :no] = Loo
    if no == n:
        return L, "block"
    Lou = solve_lower(Loo, C[:no, no:])
    L[no:, :no] = Lou.T
    try:
        Luu, _, _ = safe_chol_lower(C[no:, no:] - Lou.T @ Lou, cfg)
    except RuntimeError:
        Lf, _, _ = safe_chol_lower(C, cfg)
        return np.tril(Lf), "full"
    L[no:, no:] = Luu
    return np.tril(L), "block"


def inv_sympd(L):
    L = _f64(L)
    n = L.shape[0]
    out = np.zeros((n, n), order="F")
    lib().lko_inv_sympd(_ptr(L), n, _ptr(out))
    return out


def theta_bounds(X, y, lower_factor=0.02, upper_factor=10.0, variogram=True):
    X = _f64(X)
    y = _f64(np.asarray(y).reshape(-1))
    n, d = X.shape
    lo = np.zeros(d)
    up = np.zeros(d)
    lib().lko_theta_bounds(_ptr(X), _ptr(y), n, d, lower_factor, upper_factor, int(variogram),
                           _ptr(lo), _ptr(up))
    return lo, up
