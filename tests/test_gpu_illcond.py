"""Badly conditioned evaluations (tests/illcond_cases.py: duplicated design
points, gauss kernels at long ranges, the path of the gauss fit of the
plain-Kriging configuration, the d = 10 gauss + nugget configuration on
subsamples) against a 50-digit evaluation of the reference's formulas
(tests/golden/highprec.json, tests/golden/make_highprec.py).  The GPU's
relative error must not exceed max(1e-10 (LL) / 1e-8 (gradient), twice the
largest error of the reference algorithm under its valid FP64 orderings (the
C oracle, LAPACK's dpotrf, blocked dtrsm panels): it is never less accurate
than the reference's own arithmetic."""
import numpy as np
import pytest

from illcond_cases import arbitrate, load_truth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES, TRUTH = load_truth()


@pytest.fixture(scope="module")
def ctx():
    from libkriging_amd import _native as N
    c = N.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", sorted(TRUTH))
def test_not_less_accurate_than_the_reference_arithmetic(ctx, name):
    from libkriging_amd import _native as N
    X, y, kernel, regmodel, noise, gamma, chol = CASES[name]
    t = TRUTH[name]
    assert np.array_equal(np.asarray(t["gamma"]), gamma)
    pid = ctx.problem(X, y, O.trend_matrix(X, regmodel), kernel, noise)
    try:
        if chol:
            ctx.set_chol_config(**chol)
        ll, g, st, inc = ctx.ll_grad_batch([pid], gamma[None, :], True, N.ModelFlags(1, 1, 1.0, 0.0, 1.0))
    finally:
        ctx.set_chol_config()
        ctx.drop_problem(pid)
    P = O.OracleProblem(X, y, kernel, regmodel, noise)
    ll_o, g_o, inc_o = P.log_likelihood(gamma, cfg=O.default_cfg(**chol) if chol else None)
    e, ok = arbitrate(t, ll[0], g[0], ll_o, g_o)
    print(f"{name}: LL err GPU {e['ll_gpu']:.2e} oracle {e['ll_orc']:.2e} CPU orderings max {e['ll_cpu_max']:.2e}; "
          f"grad err GPU {e['g_gpu']:.2e} oracle {e['g_orc']:.2e} CPU orderings max {e['g_cpu_max']:.2e}; "
          f"inc {inc[0]} / {inc_o} / {t['inc']}")
    assert st[0] == 0
    assert inc[0] == inc_o == t["inc"]
    assert ok
