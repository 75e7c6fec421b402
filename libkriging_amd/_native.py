"""ctypes binding of liblkgpu.so (include/lkgpu.h).

The library must be built in-tree (python -m libkriging_amd.build); there is
no fallback: importing a GPU entry point without it raises.
"""
import ctypes
import os

import numpy as np

from ._srchash import source_hash

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblkgpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

KERNELS = {"gauss": 0, "exp": 1, "matern3_2": 2, "matern5_2": 3}
NOISES = {None: 0, "none": 0, "nugget": 1}

# exported symbols and their (restype, argtypes); every one is declared in
# include/lkgpu.h
SIGNATURES = {
    "lk_get_last_error": (ctypes.c_char_p, []),
    "lkgpu_build_id": (ctypes.c_char_p, []),
    "lkgpu_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "lkgpu_context_destroy": (None, [_vp]),
    "lkgpu_set_chol_config": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "lkgpu_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lkgpu_set_streams": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lkgpu_set_lookahead": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lkgpu_set_tuning": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "lkgpu_last_parts": (ctypes.c_int, [_vp]),
    "lkgpu_problem_create": (ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _ip]),
    "lkgpu_problem_destroy": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lkgpu_ll_grad_batch": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, ctypes.c_int, _vp, ctypes.c_int,
                                           ctypes.c_int, _dp, _dp, _ip, _ip]),
    "lkgpu_ll_grad_batch_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, ctypes.c_int, _vp, ctypes.c_int,
                                                   ctypes.c_int, _dp, _dp, _ip, _ip]),
    "lkgpu_read_item": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "lkgpu_synchronize": (ctypes.c_int, [_vp]),
    "lkgpu_set_timed_stage": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lkgpu_take_stage_time": (ctypes.c_int, [_vp, _dp, _ip]),
}

STAGES = {"build": 0, "cholesky": 1, "update0": 2, "trtri": 3, "solves": 4, "adjoint": 5}

READ_L, READ_LINV, READ_FSTAR, READ_YSTAR, READ_ESTAR, READ_X, READ_RSTAR = range(7)


class ModelFlags(ctypes.Structure):
    _fields_ = [("est_sigma2", ctypes.c_int), ("est_nugget", ctypes.c_int),
                ("sigma2", ctypes.c_double), ("nugget", ctypes.c_double), ("alpha", ctypes.c_double)]


_lib = None


def lib():
    """Load liblkgpu.so; raises if it has not been built, or was built from
    sources other than those beside it."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m libkriging_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.lkgpu_build_id.restype = ctypes.c_char_p
        built, now = L.lkgpu_build_id().decode(), source_hash()
        if now is not None and built != now:
            raise ImportError(f"{LIB_PATH} was built from other sources (build id {built}, sources {now}): "
                              "rebuild it with `python -m libkriging_amd.build`")
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def register(name, restype, argtypes):
    SIGNATURES[name] = (restype, argtypes)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().lk_get_last_error().decode())


def f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Context:
    """One lkgpu_context (one HIP device)."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().lkgpu_context_create(int(device), ctypes.byref(h)))
        self.h = h
        self._problems = {}

    def close(self):
        if self.h:
            lib().lkgpu_context_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_chol_config(self, num_nugget=1e-10, rcond_check=True, max_inc=10, min_rcond=1e-18):
        check(lib().lkgpu_set_chol_config(self.h, float(num_nugget), int(rcond_check), int(max_inc),
                                          float(min_rcond)))

    def set_graph(self, enable):
        check(lib().lkgpu_set_graph(self.h, int(enable)))

    def set_streams(self, streams):
        """most streams (1, 2, 4) a batch is split over"""
        check(lib().lkgpu_set_streams(self.h, int(streams)))

    def set_lookahead(self, enable):
        """look-ahead Cholesky for one-item evaluations of n > 8128"""
        check(lib().lkgpu_set_lookahead(self.h, int(enable)))

    def set_tuning(self, table):
        """size-bucket launch table: a dict {"buckets": [...]}, its JSON
        text, or None for the defaults"""
        import json
        txt = None if table is None else (table if isinstance(table, str) else json.dumps(table))
        check(lib().lkgpu_set_tuning(self.h, None if txt is None else txt.encode()))

    def last_parts(self):
        return lib().lkgpu_last_parts(self.h)

    def problem(self, X, y, F, kernel, noise=None):
        X = f64(X)
        if X.ndim == 1:
            X = X.reshape(-1, 1)
        y = f64(np.asarray(y).reshape(-1))
        n, d = X.shape
        F = f64(F) if F is not None else np.zeros((n, 0), order="F")
        p = F.shape[1]
        pid = ctypes.c_int()
        check(lib().lkgpu_problem_create(self.h, ptr(X), n, d, ptr(y), ptr(F) if p > 0 else None, p,
                                         KERNELS[kernel], NOISES[noise], ctypes.byref(pid)))
        self._problems[pid.value] = (n, d, p)
        return pid.value

    def drop_problem(self, pid):
        check(lib().lkgpu_problem_destroy(self.h, int(pid)))
        self._problems.pop(pid, None)

    def ll_grad_batch(self, problem_ids, gamma, return_grad=True, flags=None, noalloc=False):
        """gamma: (nb, gamma_n).  flags: ModelFlags or list of them.
        noalloc: through lkgpu_ll_grad_batch_noalloc (the device's static
        scratch array, no allocation).
        Returns (ll[nb], grad[nb, gamma_n] or None, status[nb], inc[nb])."""
        gamma = np.ascontiguousarray(np.atleast_2d(np.asarray(gamma, dtype=np.float64)))
        nb, gn = gamma.shape
        ids = np.ascontiguousarray(np.broadcast_to(np.asarray(problem_ids, dtype=np.int32), (nb,)))
        if flags is None:
            fl, fn = None, 0
        elif isinstance(flags, ModelFlags):
            fl, fn = ctypes.pointer(flags), 1
        else:
            arr = (ModelFlags * len(flags))(*flags)
            fl, fn = arr, len(flags)
        self._last_shape = [self._problems[int(i)] for i in ids]
        ll = np.zeros(nb)
        grad = np.zeros((nb, gn)) if return_grad else None
        st = np.zeros(nb, dtype=np.int32)
        inc = np.zeros(nb, dtype=np.int32)
        entry = lib().lkgpu_ll_grad_batch_noalloc if noalloc else lib().lkgpu_ll_grad_batch
        check(entry(
            self.h, nb, ids.ctypes.data_as(_ip), gamma.ctypes.data_as(_dp), gn,
            ctypes.cast(fl, _vp) if fl is not None else None, fn, int(return_grad),
            ll.ctypes.data_as(_dp), grad.ctypes.data_as(_dp) if return_grad else None,
            st.ctypes.data_as(_ip), inc.ctypes.data_as(_ip)))
        return ll, grad, st, inc

    def synchronize(self):
        check(lib().lkgpu_synchronize(self.h))

    def set_timed_stage(self, stage):
        check(lib().lkgpu_set_timed_stage(self.h, STAGES[stage] if isinstance(stage, str) else int(stage)))

    def take_stage_time(self):
        ms = ctypes.c_double()
        cnt = ctypes.c_int()
        check(lib().lkgpu_take_stage_time(self.h, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def read_item(self, item, what):
        """Host copy of a factor of batch item `item` of the last batch."""
        n, d, p = self._last_shape[item]
        if what == READ_RSTAR:
            out = np.zeros((p, p), order="F")
            check(lib().lkgpu_read_item(self.h, int(item), int(what), ptr(out)))
            return out
        cols = {READ_L: n, READ_LINV: n, READ_FSTAR: p}.get(what, 1)
        out = np.zeros((n, cols), order="F")
        check(lib().lkgpu_read_item(self.h, int(item), int(what), ptr(out)))
        return out if what in (READ_L, READ_LINV, READ_FSTAR) else out[:, 0]


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("LKGPU_DEVICE", "0")))
    return _default_ctx
