"""lkgpu_ll_grad_batch_noalloc (the allocation-free entry point: slot tables
and workspaces in the device's static scratch array) against
lkgpu_ll_grad_batch on the same inputs, at the path's bars (LL 1e-10
relative, gradient 1e-8 normwise): two calls in a row, the second with a
smaller input, then the allocating entry point again; oversize batches are
refused with a message."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _design(seed, n, d):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    return X, np.sin(3 * X).sum(axis=1) + X[:, 0]


def _compare(ctx, pids, gam):
    a = ctx.ll_grad_batch(pids, gam, True, noalloc=True)
    b = ctx.ll_grad_batch(pids, gam, True)
    for i in range(len(pids)):
        print(f"item {i}: noalloc LL {a[0][i]!r} vs {b[0][i]!r}, grad rel {rel_err(a[1][i], b[1][i]):.2e}")
        assert a[2][i] == b[2][i] == 0 and a[3][i] == b[3][i]
        assert abs(a[0][i] - b[0][i]) <= 1e-10 * abs(b[0][i])
        assert rel_err(a[1][i], b[1][i]) <= 1e-8
    return a


def test_noalloc_matches_allocating_entry_point():
    from libkriging_amd import _native as N
    ctx = N.Context(0)
    try:
        X, y = _design(1, 700, 4)
        big = ctx.problem(X, y, np.hstack([np.ones((700, 1)), X]), "matern5_2")
        X2, y2 = _design(2, 300, 3)
        small = ctx.problem(X2, y2, np.ones((300, 1)), "gauss", "nugget")
        rng = np.random.default_rng(3)
        # call 1: 6 items of n = 700 (linear trend)
        a1 = _compare(ctx, [big] * 6, np.exp(rng.uniform(np.log(0.2), 0.0, size=(6, 4))))
        # call 2, smaller: 3 items of n = 300 (nugget model, gradient over alpha)
        g2 = np.hstack([np.exp(rng.uniform(np.log(0.2), 0.0, size=(3, 3))), [[0.9], [0.95], [0.99]]])
        a2 = _compare(ctx, [small] * 3, g2)
        # the last call's factors are readable
        L = ctx.read_item(1, N.READ_L)
        assert L.shape == (300, 300) and np.all(np.triu(L, 1) == 0.0)
        # a mixed batch (two size buckets), noalloc then the allocating path
        mixed = [big, small, big]
        gm = np.array([[0.3, 0.4, 0.5, 0.6], [0.3, 0.4, 0.5, 0.9], [0.5, 0.4, 0.3, 0.2]])
        flags = [N.ModelFlags(1, 1, 1.0, 0.0, 1.0)] * 3
        a = ctx.ll_grad_batch(mixed, gm[:, :4], False, flags, noalloc=True)
        b = ctx.ll_grad_batch(mixed, gm[:, :4], False, flags)
        assert np.all(np.abs(a[0] - b[0]) <= 1e-10 * np.abs(b[0]))
        # oversize: one n = 2100 item (2 x 2112^2 doubles alone exceed 64 MiB), 65 items
        X3, y3 = _design(4, 2100, 2)
        huge = ctx.problem(X3, y3, np.ones((2100, 1)), "exp")
        with pytest.raises(RuntimeError, match="bytes of workspace"):
            ctx.ll_grad_batch([huge], np.array([[0.3, 0.3]]), True, noalloc=True)
        with pytest.raises(RuntimeError, match="at most 64 items"):
            ctx.ll_grad_batch([small] * 65, np.tile([0.3, 0.3, 0.3, 0.9], (65, 1)), False, noalloc=True)
        # the allocating entry point is unaffected by the refusals
        ll, _, st, _ = ctx.ll_grad_batch([huge], np.array([[0.3, 0.3]]), False)
        assert st[0] == 0 and np.isfinite(ll[0])
        assert a1 is not None and a2 is not None
    finally:
        ctx.close()
