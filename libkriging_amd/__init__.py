"""libkriging_amd: the Kriging log-likelihood hot path of libKriging
(covariance build, Cholesky with jitter retry, triangular solves / GLS trend,
concentrated LL and its adjoint gradient inside the L-BFGS-B fit), written
for one MI355X (gfx950) in FP64 HIP behind a C ABI (include/lkgpu.h).

Python surface mirrors pylibkriging: ``Kriging`` (= lk.Kriging),
``NestedKriging`` (= lk.NestedKriging, plain submodels), ``Optim``;
plus ``Context`` (the batched low-level API) and, when torch is importable,
``KrigingLogLikelihood`` (a torch.autograd.Function whose backward is the
adjoint gradient).
"""
from ._native import Context, ModelFlags  # noqa: F401
from .kriging import Kriging, Optim, lbfgsb_minimize, set_num_nugget, check_chol_rcond, reset_seed  # noqa: F401
from .nested import NestedKriging  # noqa: F401
from . import linalg  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    if name in ("KrigingLogLikelihood", "log_likelihood"):
        from . import torch_ll
        return getattr(torch_ll, name)
    raise AttributeError(name)
