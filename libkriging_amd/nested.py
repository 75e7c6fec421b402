"""Python surface of NestedKriging, mirroring pylibkriging's ``lk.NestedKriging``
(= WrappedPyNestedKriging, bindings/Python/pylibkriging/src/_pylibkriging/
pylibkriging.cpp:224-285, NestedKriging_binding.cpp): same constructor forms,
keyword names, defaults and return tuples, over the C ABI of liblkgpu.so
(include/lkgpu.h, lk_nested_kriging_*).  Plain Kriging submodels only:
``warping`` must be empty.

Extensions: ``fit_groups`` (caller-given, possibly uneven groups),
``submodel(i)`` (a view of group i's Kriging model) and ``predict_fallbacks``.
"""
import ctypes

import numpy as np

from . import _native as N
from .kriging import Kriging, _b, _col, _mat

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p
_i = ctypes.c_int

N.register("lk_nested_kriging_new", _vp, [_cp, _cp, _cp, _i])
N.register("lk_nested_kriging_new_fit", _vp, [_dp, _i, _dp, _i, _i, _cp, _i, _cp, _cp, _i, ctypes.c_void_p, _i, _cp,
                                              _cp, _cp, _dp, _i, _dp, _i, _i, _dp, _i, _i])
N.register("lk_nested_kriging_delete", None, [_vp])
N.register("lk_nested_kriging_fit", _i, [_vp, _dp, _i, _dp, _i, _i, _i, ctypes.c_void_p, _i, _cp, _cp, _cp, _dp, _i,
                                         _dp, _i, _i, _dp, _i, _i])
N.register("lk_nested_kriging_fit_groups", _i, [_vp, _dp, _i, _dp, _i, _i, _ip, _i, _cp, _cp, _cp, _dp, _i, _dp, _i,
                                                _i, _dp, _i, _i])
N.register("lk_nested_kriging_predict", _i, [_vp, _dp, _i, _i, _i, _dp, _dp])
for _name in ("lk_nested_kriging_summary", "lk_nested_kriging_kernel", "lk_nested_kriging_aggregation"):
    N.register(_name, _cp, [_vp])
N.register("lk_nested_kriging_nb_groups", _i, [_vp])
N.register("lk_nested_kriging_get_theta", _i, [_vp, _dp, _ip])
N.register("lk_nested_kriging_get_sigma2", ctypes.c_double, [_vp])
N.register("lk_nested_kriging_get_beta0", ctypes.c_double, [_vp])
N.register("lk_nested_kriging_groups", _i, [_vp, _ip])
N.register("lk_nested_kriging_submodel", _vp, [_vp, _i])
N.register("lk_nested_kriging_set_predict_chunk", _i, [_vp, _i])
N.register("lk_nested_kriging_predict_fallbacks", _i, [_vp])
N.register("lk_nested_kriging_fit_statistics", _i, [_vp, _ip, _ip])
N.register("lk_nested_kriging_group_fits", _i, [_vp, _dp, _dp, _dp])
N.register("lk_nested_kriging_time_pairs", _i, [_vp, _i])
N.register("lk_nested_kriging_pair_time", _i, [_vp, _dp, _ip, _ip])


def _nparams(parameters, d):
    """parameters dict (sigma2, is_sigma2_estim, theta, is_theta_estim, beta,
    is_beta_estim) -> C arguments; theta is one row (libkriging_c.cpp:1675-1676)."""
    p = dict(parameters or {})
    sig = p.get("sigma2")
    th = p.get("theta")
    be = p.get("beta")
    sig_c = ctypes.byref(ctypes.c_double(float(sig))) if sig is not None else None
    th_a = np.ascontiguousarray(np.asarray(th, dtype=np.float64).reshape(-1)) if th is not None else None
    be_a = _col(be) if be is not None else None
    return (sig_c, int(p.get("is_sigma2_estim", True)),
            th_a, int(p.get("is_theta_estim", True)),
            be_a, int(p.get("is_beta_estim", True)))


def _ptr_or_none(a):
    return None if a is None else N.ptr(a)


class NestedKriging:
    """NestedKriging(kernel) |
    NestedKriging(y, X, kernel, nb_groups, aggregation="NK", partition="kmeans",
                  seed=123, regmodel="constant", optim="BFGS", objective="LL",
                  parameters={}, warping=[])"""

    def __init__(self, *args, **kwargs):
        self._h = None
        names = ["y", "X", "kernel", "nb_groups", "aggregation", "partition", "seed", "regmodel", "optim",
                 "objective", "parameters", "warping"]
        if len(args) + len(kwargs) == 1 and (args and isinstance(args[0], str) or "kernel" in kwargs):
            kernel = kwargs.get("kernel", args[0] if args else None)
            self._h = N.lib().lk_nested_kriging_new(_b(kernel), b"NK", b"kmeans", 123)
            if not self._h:
                raise ValueError(N.lib().lk_get_last_error().decode())
            return
        vals = dict(zip(names, args))
        vals.update(kwargs)
        if vals.get("warping"):
            raise ValueError("warping (WarpKriging submodels) is not supported on this path")
        self._h = N.lib().lk_nested_kriging_new(_b(vals["kernel"]), _b(vals.get("aggregation", "NK")),
                                                _b(vals.get("partition", "kmeans")), int(vals.get("seed", 123)))
        if not self._h:
            raise ValueError(N.lib().lk_get_last_error().decode())
        self.fit(vals["y"], vals["X"], vals["nb_groups"], vals.get("regmodel", "constant"), vals.get("optim", "BFGS"),
                 vals.get("objective", "LL"), vals.get("parameters", None), vals.get("warping", None))

    def __del__(self):
        try:
            if self._h:
                N.lib().lk_nested_kriging_delete(self._h)
                self._h = None
        except Exception:
            pass

    def fit(self, y, X, nb_groups, regmodel="constant", optim="BFGS", objective="LL", parameters=None, warping=None):
        if warping:
            raise ValueError("warping (WarpKriging submodels) is not supported on this path")
        y = _col(y)
        X = _mat(X)
        n, d = X.shape
        sig, sig_e, th, th_e, be, be_e = _nparams(parameters, d)
        N.check(N.lib().lk_nested_kriging_fit(
            self._h, N.ptr(y), y.size, N.ptr(X), n, d, int(nb_groups), None, 0, _b(regmodel), _b(optim),
            _b(objective), sig, sig_e, _ptr_or_none(th), 0 if th is None else th.size, th_e, _ptr_or_none(be),
            0 if be is None else be.size, be_e))

    def fit_groups(self, y, X, group_of, regmodel="constant", optim="BFGS", objective="LL", parameters=None):
        """fit on given groups: group_of[i] in 0..G-1 is the group of point i"""
        y = _col(y)
        X = _mat(X)
        n, d = X.shape
        g = np.ascontiguousarray(np.asarray(group_of, dtype=np.int32).reshape(-1))
        if g.size != n:
            raise ValueError("group_of must have one entry per point")
        sig, sig_e, th, th_e, be, be_e = _nparams(parameters, d)
        N.check(N.lib().lk_nested_kriging_fit_groups(
            self._h, N.ptr(y), y.size, N.ptr(X), n, d, g.ctypes.data_as(_ip), int(g.max()) + 1 if n else 0,
            _b(regmodel), _b(optim), _b(objective), sig, sig_e, _ptr_or_none(th), 0 if th is None else th.size,
            th_e, _ptr_or_none(be), 0 if be is None else be.size, be_e))

    def predict(self, X, return_stdev=True):
        X = _mat(X)
        m, d = X.shape
        mean = np.empty(m)
        sd = np.empty(m)
        N.check(N.lib().lk_nested_kriging_predict(self._h, N.ptr(X), m, d, int(bool(return_stdev)), N.ptr(mean),
                                                  N.ptr(sd)))
        return mean, (sd if return_stdev else np.empty(0))

    def summary(self):
        return N.lib().lk_nested_kriging_summary(self._h).decode()

    def __repr__(self):
        return self.summary()

    def kernel(self):
        return N.lib().lk_nested_kriging_kernel(self._h).decode()

    def warping(self):
        return []

    def aggregation(self):
        return N.lib().lk_nested_kriging_aggregation(self._h).decode()

    def nb_groups(self):
        return N.lib().lk_nested_kriging_nb_groups(self._h)

    def groups(self):
        n = sum(self.submodel(k).y().size for k in range(self.nb_groups()))
        g = np.empty(n, dtype=np.int32)
        N.check(N.lib().lk_nested_kriging_groups(self._h, g.ctypes.data_as(_ip)))
        return [np.flatnonzero(g == k).astype(np.uint64) for k in range(self.nb_groups())]

    def theta(self):
        n = ctypes.c_int(0)
        N.check(N.lib().lk_nested_kriging_get_theta(self._h, None, ctypes.byref(n)))
        out = np.empty(n.value)
        N.check(N.lib().lk_nested_kriging_get_theta(self._h, N.ptr(out), ctypes.byref(n)))
        return out

    def sigma2(self):
        return N.lib().lk_nested_kriging_get_sigma2(self._h)

    def beta0(self):
        return N.lib().lk_nested_kriging_get_beta0(self._h)

    def submodel(self, i):
        h = N.lib().lk_nested_kriging_submodel(self._h, int(i))
        if not h:
            raise IndexError(N.lib().lk_get_last_error().decode())
        return Kriging._borrowed(h, self)

    def X(self):
        # the design in the original order, from the submodels and groups
        parts = [(np.asarray(g, dtype=np.int64), self.submodel(k).X()) for k, g in enumerate(self.groups())]
        n = sum(len(g) for g, _ in parts)
        d = parts[0][1].shape[1]
        out = np.empty((n, d))
        for g, x in parts:
            out[g] = x
        return out

    def y(self):
        parts = [(np.asarray(g, dtype=np.int64), self.submodel(k).y()) for k, g in enumerate(self.groups())]
        out = np.empty(sum(len(g) for g, _ in parts))
        for g, v in parts:
            out[g] = v
        return out

    def set_predict_chunk(self, chunk):
        N.check(N.lib().lk_nested_kriging_set_predict_chunk(self._h, int(chunk)))

    def set_warp_subsample(self, m):
        raise ValueError("warping (WarpKriging submodels) is not supported on this path")

    def fit_statistics(self):
        """(evaluations, batched pipelines) of the groups' optimisation in the last fit"""
        e, b = ctypes.c_int(), ctypes.c_int()
        N.check(N.lib().lk_nested_kriging_fit_statistics(self._h, ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def group_fits(self):
        """(theta (nb_groups x d), sigma2, beta0) each group's fit reached
        before the hyper-parameters were unified"""
        p, d = self.nb_groups(), self.theta().size
        th, s2, b0 = np.empty((p, d)), np.empty(p), np.empty(p)
        N.check(N.lib().lk_nested_kriging_group_fits(self._h, N.ptr(th), N.ptr(s2), N.ptr(b0)))
        return th, s2, b0

    def time_pairs(self, enable=True):
        """HIP-event timing of the pair-contraction launches of later predicts"""
        N.check(N.lib().lk_nested_kriging_time_pairs(self._h, int(bool(enable))))

    def pair_time(self):
        """(ms, launches, points) of the timed pair launches since the last call"""
        ms, nl, npt = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        N.check(N.lib().lk_nested_kriging_pair_time(self._h, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(npt)))
        return ms.value, nl.value, npt.value

    def predict_fallbacks(self):
        return N.lib().lk_nested_kriging_predict_fallbacks(self._h)
