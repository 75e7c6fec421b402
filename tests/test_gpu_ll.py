"""GPU parity of the batched LL + adjoint gradient (through the C ABI,
lkgpu_ll_grad_batch) against the CPU oracle and the reference's golden
vectors.

Tolerances (the stated FP64 parity bar of this path): log-likelihood within
1e-10 relative, gradient within 1e-8 normwise relative.  Badly conditioned
evaluations (cond(R) 1e7 .. 1e14, tests/illcond_cases.py) are arbitrated by a
50-digit evaluation of the same formulas instead: there the GPU's error must
not exceed twice the oracle's (the reference's substitution arithmetic).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err
from illcond_cases import arbitrate, load_truth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LL_TOL = 1e-10
GRAD_TOL = 1e-8


@pytest.fixture(scope="module")
def ctx():
    from libkriging_amd import _native as N
    c = N.Context(0)
    yield c
    c.close()


def _design(seed, n, d, kernel, regmodel="constant", noise=None):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sin(2 * np.pi * X).sum(axis=1) + 0.5 * X[:, 0] * X[:, -1]
    return X, y


def _flags(noise, est_sigma2=True, est_nugget=True, sigma2=1.0, nugget=0.0, alpha=1.0):
    from libkriging_amd import _native as N
    return N.ModelFlags(int(est_sigma2), int(est_nugget), sigma2, nugget, alpha)


_CASES, _TRUTH = load_truth()


def _check(ctx, X, y, kernel, regmodel, noise, gammas, flags_kw=None, names=None):
    """GPU against the oracle at the plain bars; a slot named in `names` that
    has a 50-digit value is arbitrated by it instead"""
    flags_kw = flags_kw or {}
    F = O.trend_matrix(X, regmodel)
    pid = ctx.problem(X, y, F, kernel, noise)
    try:
        gam = np.atleast_2d(gammas)
        ll, g, st, inc = ctx.ll_grad_batch([pid] * gam.shape[0], gam, True, _flags(noise, **flags_kw))
        P = O.OracleProblem(X, y, kernel, regmodel, noise, **{k: v for k, v in flags_kw.items()})
        for b in range(gam.shape[0]):
            ll_r, g_r, inc_r = P.log_likelihood(gam[b])
            name = names[b] if names else None
            assert st[b] == 0
            assert inc[b] == inc_r
            if name in _TRUTH:
                Xc, _, _, _, _, gc, _ = _CASES[name]
                assert np.array_equal(Xc, X) and np.array_equal(gc, gam[b])
                assert inc[b] == _TRUTH[name]["inc"]
                e, ok = arbitrate(_TRUTH[name], ll[b], g[b], ll_r, g_r)
                print(f"slot {b} ({name}): vs 50 digits LL {e['ll_gpu']:.2e} (oracle {e['ll_orc']:.2e}, CPU "
                      f"orderings max {e['ll_cpu_max']:.2e}), grad {e['g_gpu']:.2e} (oracle {e['g_orc']:.2e}, "
                      f"max {e['g_cpu_max']:.2e}), inc {inc[b]}")
                assert ok
                continue
            print(f"slot {b}: ll {ll[b]!r} vs {ll_r!r} (rel {abs(ll[b] - ll_r) / abs(ll_r):.2e}), "
                  f"grad rel {rel_err(g[b], g_r):.2e}, inc {inc[b]} vs {inc_r}")
            assert abs(ll[b] - ll_r) <= LL_TOL * abs(ll_r)
            assert rel_err(g[b], g_r) <= GRAD_TOL
    finally:
        ctx.drop_problem(pid)


def _load(prefix):
    X = np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-X.csv"), delimiter=",")
    y = np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-y.csv"), delimiter=",")
    llr = float(np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-result-logLikelihood.csv"), delimiter=","))
    gr = np.atleast_1d(np.genfromtxt(os.path.join(GOLDEN, f"{prefix}-result-logLikelihoodGrad.csv"), delimiter=","))
    return X.reshape(len(y), -1), y, llr, gr


@pytest.mark.parametrize("prefix", ["data1-scal"] + [f"data2-grad-{i}" for i in range(1, 11)])
def test_golden_vectors(ctx, prefix):
    X, y, llr, gr = _load(prefix)
    pid = ctx.problem(X, y, np.ones((len(y), 1)), "gauss")
    try:
        ll, g, st, inc = ctx.ll_grad_batch([pid], 0.3 * np.ones((1, X.shape[1])), True)
    finally:
        ctx.drop_problem(pid)
    print(prefix, ll[0], llr, g[0], gr)
    assert st[0] == 0
    assert abs(ll[0] - llr) <= LL_TOL * abs(llr)
    assert rel_err(g[0], gr) <= GRAD_TOL


@pytest.mark.parametrize("kernel", ["gauss", "exp", "matern3_2", "matern5_2"])
@pytest.mark.parametrize("n,d", [(37, 2), (64, 3), (130, 6), (300, 4)])
def test_kernels_constant(ctx, kernel, n, d):
    X, y = _design(n * 7 + d, n, d, kernel)
    rng = np.random.default_rng(n + d)
    gam = np.exp(rng.uniform(np.log(0.2), np.log(1.0), size=(3, d)))
    _check(ctx, X, y, kernel, "constant", None, gam, names=[f"{kernel}{n}_{b}" for b in range(3)])


@pytest.mark.parametrize("kernel", ["gauss", "matern5_2"])
def test_linear_trend(ctx, kernel):
    X, y = _design(5, 200, 5, kernel)
    gam = np.full((2, 5), 0.5)
    gam[1] = [0.3, 0.6, 0.9, 0.4, 0.7]
    _check(ctx, X, y + X @ np.arange(1, 6), kernel, "linear", None, gam)


@pytest.mark.parametrize("kernel", ["gauss", "exp", "matern3_2", "matern5_2"])
def test_nugget_alpha_gradient(ctx, kernel):
    X, y = _design(11, 150, 3, kernel)
    y = y + np.random.default_rng(2).normal(scale=0.05, size=y.shape)
    gam = np.array([[0.4, 0.5, 0.6, 0.9], [0.2, 0.8, 0.3, 0.99]])
    _check(ctx, X, y, kernel, "constant", "nugget", gam)


def test_nugget_fixed_variants(ctx):
    X, y = _design(3, 120, 2, "gauss")
    gam = np.array([[0.4, 0.5, 0.9]])
    _check(ctx, X, y, "gauss", "constant", "nugget", gam, dict(est_nugget=False, nugget=0.01))
    _check(ctx, X, y, "gauss", "constant", "nugget", gam[:, :2],
           dict(est_sigma2=False, est_nugget=False, sigma2=1.5, nugget=0.02, alpha=0.9))


def test_fixed_sigma2(ctx):
    X, y = _design(4, 90, 3, "matern3_2")
    _check(ctx, X, y, "matern3_2", "constant", None, np.array([[0.3, 0.4, 0.5]]),
           dict(est_sigma2=False, sigma2=2.0))


def test_jitter_ladder_duplicate_points(ctx):
    # duplicated design rows make R singular: the retry ladder must add the
    # same cumulative jitter as LinearAlgebra.cpp:68-98 and give the same LL
    X, y = _design(9, 80, 2, "gauss")
    X = np.vstack([X, X[:5]])
    y = np.concatenate([y, y[:5]])
    _check(ctx, X, y, "gauss", "constant", None, np.array([[0.3, 0.3]]), names=["duplicates"])


def test_uneven_batch(ctx):
    # designs of different sizes in the same launches
    pids, refs, gams = [], [], []
    for i, n in enumerate([70, 300, 129, 64, 500]):
        X, y = _design(100 + i, n, 3, "matern3_2")
        pids.append(ctx.problem(X, y, np.ones((n, 1)), "matern3_2"))
        refs.append(O.OracleProblem(X, y, "matern3_2"))
        gams.append([0.3 + 0.05 * i, 0.5, 0.4])
    try:
        ll, g, st, inc = ctx.ll_grad_batch(pids, np.array(gams), True)
        for b in range(len(pids)):
            ll_r, g_r, _ = refs[b].log_likelihood(gams[b])
            assert st[b] == 0
            assert abs(ll[b] - ll_r) <= LL_TOL * abs(ll_r)
            assert rel_err(g[b], g_r) <= GRAD_TOL
    finally:
        for p in pids:
            ctx.drop_problem(p)


def test_graph_replay_matches_eager(ctx):
    # several replays of the captured pipeline, with eager runs in between,
    # must all equal the eager result bit for bit
    X, y = _design(21, 257, 4, "exp")
    pid = ctx.problem(X, y, np.ones((257, 1)), "exp")
    try:
        gam = np.exp(np.random.default_rng(0).uniform(-1.5, 0.5, size=(6, 4)))
        ctx.set_graph(False)
        ll0, g0, _, _ = ctx.ll_grad_batch([pid] * 6, gam, True)
        ctx.set_graph(True)
        for rep in range(3):
            ll1, g1, _, _ = ctx.ll_grad_batch([pid] * 6, gam, True)
            assert np.array_equal(ll0, ll1)
            assert np.array_equal(g0, g1)
            ctx.set_graph(False)
            ll2, g2, _, _ = ctx.ll_grad_batch([pid] * 6, gam[::-1], True)
            assert np.array_equal(ll2, ll0[::-1])
            ctx.set_graph(True)
    finally:
        ctx.set_graph(True)
        ctx.drop_problem(pid)


def test_stream_split_matches_one_stream(ctx):
    # a batch split over 4 or 2 streams (4 or 2 parts, each on its own graph
    # branch) equals the one-stream eager result bit for bit, on every replay,
    # with eager runs in between; uniform and uneven batches
    pids = []
    try:
        X, y = _design(41, 200, 3, "matern5_2")
        pids.append(ctx.problem(X, y, np.ones((200, 1)), "matern5_2"))
        uneven = []
        for i, n in enumerate([70, 300, 129, 64, 250, 33, 190, 96] * 4):
            X, y = _design(200 + i, n, 3, "matern5_2")
            uneven.append(ctx.problem(X, y, np.ones((n, 1)), "matern5_2"))
        pids += uneven
        rng = np.random.default_rng(2)
        for ids in ([pids[0]] * 64, uneven):
            gam = np.exp(rng.uniform(-1.5, 0.5, size=(len(ids), 3)))
            ctx.set_streams(1)
            ctx.set_graph(False)
            ll0, g0, st0, inc0 = ctx.ll_grad_batch(ids, gam, True)
            assert (st0 == 0).all()
            for S in (4, 2):
                ctx.set_streams(S)
                for rep in range(3):
                    ctx.set_graph(True)
                    ll1, g1, st1, inc1 = ctx.ll_grad_batch(ids, gam, True)
                    assert np.array_equal(ll1, ll0) and np.array_equal(g1, g0)
                    assert np.array_equal(st1, st0) and np.array_equal(inc1, inc0)
                    ctx.set_graph(False)
                    ll2, g2, _, _ = ctx.ll_grad_batch(ids, gam, True)
                    assert np.array_equal(ll2, ll0) and np.array_equal(g2, g0)
    finally:
        ctx.set_streams(4)
        ctx.set_graph(True)
        for p in pids:
            ctx.drop_problem(p)


def test_batch_equals_single(ctx):
    # a batched evaluation must equal the same evaluations one per launch
    X, y = _design(31, 400, 6, "exp")
    pid = ctx.problem(X, y, np.ones((400, 1)), "exp")
    try:
        gam = np.exp(np.random.default_rng(1).uniform(np.log(0.05), np.log(2.0), size=(8, 6)))
        llb, gb, _, _ = ctx.ll_grad_batch([pid] * 8, gam, True)
        for b in range(8):
            l1, g1, _, _ = ctx.ll_grad_batch([pid], gam[b:b + 1], True)
            assert l1[0] == llb[b]
            assert np.array_equal(g1[0], gb[b])
    finally:
        ctx.drop_problem(pid)


def test_bucket_parts_match_one_stream(ctx):
    """an uneven batch spanning several size buckets runs as one part per
    bucket (one stream each, launches sized to the part's largest matrix)
    under any tuning table, and gives the one-stream results bit for bit;
    items keep their request order whatever the slot order"""
    # 1 .. 18 blocks: default buckets <= 4, <= 8, <= 16, <= 32 (four parts);
    # the second table's <= 4, <= 16 and the rest (three parts)
    sizes = [70, 600, 129, 64, 1100, 33, 190, 96, 530, 260]
    pids = []
    try:
        for i, n in enumerate(sizes):
            X, y = _design(500 + i, n, 3, "matern3_2")
            pids.append(ctx.problem(X, y, np.ones((n, 1)), "matern3_2"))
        gam = np.exp(np.random.default_rng(5).uniform(-1.5, 0.0, size=(len(sizes), 3)))
        ctx.set_streams(1)
        ll0, g0, st0, inc0 = ctx.ll_grad_batch(pids, gam, True)
        assert ctx.last_parts() == 1 and (st0 == 0).all()
        ctx.set_streams(4)
        for table, parts in ((None, 4), ({"buckets": [{"max_nblk": 4, "panel_blocks": 1, "streams": 1},
                                                      {"max_nblk": 16, "panel_blocks": 2, "streams": 2},
                                                      {"max_nblk": None, "panel_blocks": 8, "streams": 4}]}, 3)):
            ctx.set_tuning(table)
            for rep in range(2):
                ll1, g1, st1, inc1 = ctx.ll_grad_batch(pids, gam, True)
                assert ctx.last_parts() == parts
                assert np.array_equal(ll1, ll0) and np.array_equal(g1, g0)
                assert np.array_equal(st1, st0) and np.array_equal(inc1, inc0)
        # request order preserved: the oracle on a few items
        for b in (1, 4, 5):
            X, y = _design(500 + b, sizes[b], 3, "matern3_2")
            ll_r, g_r, _ = O.OracleProblem(X, y, "matern3_2").log_likelihood(gam[b])
            assert abs(ll0[b] - ll_r) <= LL_TOL * abs(ll_r)
            assert rel_err(g0[b], g_r) <= GRAD_TOL
    finally:
        ctx.set_tuning(None)
        ctx.set_streams(4)
        for p in pids:
            ctx.drop_problem(p)


def test_lookahead_matches_plain(ctx):
    """the look-ahead Cholesky of a one-item evaluation of 128 blocks (the
    next panel's chain beside the bulk update on a CU-masked stream, issued
    without a graph) gives the plain pipeline's results bit for bit: graph
    and eager runs without it, repeated runs and a timed-stage run with it"""
    n, d = 8192, 10
    rng = np.random.default_rng(125)
    X = rng.uniform(size=(n, d))
    y = rng.normal(size=n)
    theta = np.exp(rng.uniform(np.log(0.3), np.log(1.5), size=d))
    gam = np.concatenate([theta, [1.0 / (1.0 + 1e-6)]])[None, :]
    pid = ctx.problem(X, y, np.ones((n, 1)), "gauss", "nugget")
    try:
        ctx.set_lookahead(False)
        ll0, g0, st0, inc0 = ctx.ll_grad_batch([pid], gam, True)
        assert st0[0] == 0
        ctx.set_graph(False)
        ll1, g1, _, _ = ctx.ll_grad_batch([pid], gam, True)
        assert np.array_equal(ll1, ll0) and np.array_equal(g1, g0)
        ctx.set_graph(True)
        ctx.set_lookahead(True)
        for rep in range(2):
            ll2, g2, st2, inc2 = ctx.ll_grad_batch([pid], gam, True)
            assert np.array_equal(ll2, ll0) and np.array_equal(g2, g0)
            assert st2[0] == 0 and inc2[0] == inc0[0]
        ctx.set_timed_stage("cholesky")
        ll3, g3, _, _ = ctx.ll_grad_batch([pid], gam, True)
        ms, count = ctx.take_stage_time()
        assert count == 1 and ms > 0.0
        assert np.array_equal(ll3, ll0) and np.array_equal(g3, g0)
    finally:
        ctx.set_timed_stage(-1)
        ctx.set_lookahead(True)
        ctx.set_graph(True)
        ctx.drop_problem(pid)
