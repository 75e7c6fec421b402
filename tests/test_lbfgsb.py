"""The fit's L-BFGS-B (host code, lk_lbfgsb_minimize) against scipy's
fmin_l_bfgs_b, which wraps the Fortran L-BFGS-B 3.0 the reference's
lbfgsb_cpp ports: same iterates, evaluation counts and stopping reasons on
smooth test problems with active and inactive bounds."""
import numpy as np
import pytest
from scipy.optimize import fmin_l_bfgs_b


def rosen(x):
    f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] = -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return f, g


def quad(x):
    A = np.diag([1.0, 10.0, 100.0, 3.0])
    b = np.array([5.0, -3.0, 1.0, 0.2])
    return 0.5 * x @ A @ x - b @ x, A @ x - b


CASES = [
    (rosen, [-1.2, 1.0], [-2, -2], [2, 2]),
    (rosen, [0.1] * 5, [-2] * 5, [0.8] * 5),
    (rosen, [2.5, 0.6, 2, 1], [0.5] * 4, [3] * 4),
    (quad, [0, 0, 0, 0], [-1] * 4, [1] * 4),
    (quad, [0.9, -0.9, 0.5, 0.0], [0.5, -2, -2, -2], [2, 2, 0.001, 2]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("max_iter,pgtol,factr", [(20, 1e-3, 1e10), (200, 1e-5, 1e7)])
def test_matches_scipy(case, max_iter, pgtol, factr):
    from libkriging_amd import lbfgsb_minimize
    fun, x0, lo, up = CASES[case]
    x, f, it, ev, task = lbfgsb_minimize(fun, x0, lo, up, max_iter=max_iter, pgtol=pgtol, factr=factr, m=5)
    xs, fs, info = fmin_l_bfgs_b(fun, np.array(x0, float), bounds=list(zip(lo, up)), m=5, factr=factr,
                                 pgtol=pgtol, maxiter=max_iter)
    stask = info["task"] if isinstance(info["task"], str) else info["task"].decode()
    assert it == info["nit"]
    assert ev == info["funcalls"]
    assert abs(f - fs) <= 1e-9 * max(1.0, abs(fs))
    assert np.allclose(x, xs, rtol=1e-7, atol=1e-9)
    key = {"STOP": "STOP", "CONVERGENCE: NORM": "NORM", "CONVERGENCE: REL": "REL"}
    for a, b in [("STOP", "STOP"), ("CONVERGENCE: NORM", "CONVERGENCE: NORM"), ("CONVERGENCE: REL", "CONVERGENCE: REL")]:
        if stask.startswith(b):
            assert task.startswith(a), (task, stask)
    del key


def test_start_outside_box_is_projected():
    from libkriging_amd import lbfgsb_minimize
    x, f, it, ev, task = lbfgsb_minimize(quad, [5, 5, 5, 5], [-1] * 4, [1] * 4, max_iter=50, pgtol=1e-8, factr=10)
    assert np.all(x <= 1) and np.all(x >= -1)
    xs, fs, _ = fmin_l_bfgs_b(quad, np.array([5.0] * 4), bounds=[(-1, 1)] * 4, m=5, factr=10, pgtol=1e-8, maxiter=50)
    assert np.allclose(x, xs, atol=1e-8)
