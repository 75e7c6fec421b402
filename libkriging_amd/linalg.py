"""The public LinearAlgebra routines of libKriging's R binding
(bindings/R/rlibkriging/src/linalg.cpp:10-71), on the device: safe Cholesky
with the jitter retry ladder, the condition estimates of a factor, the block
Cholesky update, and the process-wide ladder settings."""
from .kriging import (  # noqa: F401
    check_chol_rcond,
    chol_block,
    chol_rcond_checked,
    chol_safe,
    get_num_nugget,
    rcond_approx_chol,
    rcond_chol,
    set_num_nugget,
)
