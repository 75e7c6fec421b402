"""The public LinearAlgebra routines on the device against the oracle:

* safe_chol_lower (LinearAlgebra.cpp:43-98) on well-conditioned matrices
  and on the edge cases of tests/LinearAlgebraTest.cpp:462-655 (nearly
  singular, a small negative eigenvalue, eigenvalues 1e-14..1, a rank-3
  matrix plus 1e-13, an all-but-collinear correlation matrix, eigenvalues
  1e-16..1, symmetrised noise), rebuilt with numpy's generator: the jitter
  steps equal the oracle's, the factor reproduces M + jitter to backward-
  error level and, where M is well conditioned, equals the oracle's factor;
* rcond_chol (LinearAlgebra.cpp:106-115) against LAPACK dtrcon squared;
* chol_block (LinearAlgebra.cpp:257-299) on random and covariance
  matrices, old blocks of 1..n-1 rows (inside and across 64-row blocks),
  and its fall-back to the full factor.
"""
import numpy as np
import pytest
from scipy.linalg import lapack

from illcond_cases import blocked_cholesky
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def la():
    import libkriging_amd as lk
    yield lk.linalg
    lk.linalg.check_chol_rcond(True)
    lk.linalg.set_num_nugget(1e-10)


def spd(n, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, -np.log10(cond), n)
    return (Q * ev) @ Q.T


def with_eigs(n, seed, eigs):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.asarray(eigs, dtype=float)) @ Q.T


def edge_cases():
    rng = np.random.default_rng(7)
    out = {}
    ev = np.linspace(0.5, 2.0, 10)
    ev[0] = 1e-12
    out["nearly_singular"] = with_eigs(10, 42, ev)
    ev = np.linspace(0.5, 2.0, 8)
    ev[0] = -1e-14
    out["small_negative_eig"] = with_eigs(8, 123, ev)
    out["ill_conditioned"] = with_eigs(10, 456, np.linspace(1e-14, 1.0, 10))
    A = rng.standard_normal((10, 3))
    M = A @ A.T
    M[np.diag_indices(10)] += 1e-13
    out["rank_deficient"] = M
    M = np.full((5, 5), 1.0 - 1e-20)
    M[np.diag_indices(5)] = 1.0
    out["collinear_correlation"] = M
    out["tiny_eigenvalues"] = with_eigs(10, 789, [10.0 ** (-16.0 + 1.6 * i) for i in range(10)])
    M = spd(8, 999, cond=10.0)
    M = M + rng.standard_normal((8, 8)) * 1e-14
    out["asymmetric_noise"] = (M + M.T) / 2.0
    return out


def jittered(M, inc, num_nugget=1e-10):
    J = M.copy()
    for k in range(inc):
        J[np.diag_indices(M.shape[0])] += num_nugget * 10.0 ** k
    return J


@pytest.mark.parametrize("n", [1, 5, 64, 65, 200])
def test_safe_chol_well_conditioned(la, n):
    M = spd(n, n, cond=1e4)
    L, inc = la.chol_safe(M, return_inc=True)
    Lo, inco, _ = O.safe_chol_lower(M)
    assert inc == inco == 0
    assert np.array_equal(L, np.tril(L))
    err = np.max(np.abs(L - Lo)) / np.max(np.abs(Lo))
    print(f"safe_chol n={n}: rel err vs oracle {err:.2e}")
    assert err < 1e-12


def ladder_inc(M, chol):
    """Jitter steps of the ladder (LinearAlgebra.cpp:68-98) when its
    factorisation is `chol` (None = the oracle's)."""
    if chol is None:
        return O.safe_chol_lower(M)[1]
    for k in range(11):
        try:
            L = chol(jittered(M, k))
        except np.linalg.LinAlgError:
            continue
        if np.all(np.isfinite(L)) and np.all(np.diag(L) > 0) and _dtrcon2(L) >= 1e-18:
            return k
    return None


def _blocked(nb):
    def f(A):
        L = blocked_cholesky(A, nb)
        if not np.all(np.isfinite(L)):
            raise np.linalg.LinAlgError("not positive definite")
        return L
    return f


@pytest.mark.parametrize("name", sorted(edge_cases()))
def test_safe_chol_edge_cases(la, name):
    # the jitter steps must equal the oracle's; where the decision sits at
    # rounding level (valid FP64 orderings of the factorisation disagree on
    # it: LAPACK dpotrf, blocked right-looking with 64 / 32 columns), they
    # must be one of those orderings' decisions
    M = edge_cases()[name]
    Lo, inco, _ = O.safe_chol_lower(M)
    L, inc = la.chol_safe(M, return_inc=True)
    n = M.shape[0]
    allowed = {inco} | {ladder_inc(M, c) for c in (np.linalg.cholesky, _blocked(64), _blocked(32))}
    back = np.max(np.abs(L @ L.T - jittered(M, inc))) / np.max(np.abs(M))
    print(f"{name}: inc {inc} (oracle {inco}, CPU orderings {sorted(allowed)}), backward error {back:.2e}")
    assert inc == inco if len(allowed) == 1 else inc in allowed
    assert back < 1e-14 * n
    assert np.array_equal(L, np.tril(L))


def test_safe_chol_failures(la):
    # -2 I: the ladder's total jitter (1e-10 (1 + 10 + ... + 1e10) ~ 1.11)
    # cannot make it positive definite
    M = -2.0 * np.eye(4)
    with pytest.raises(RuntimeError):
        O.safe_chol_lower(M)
    with pytest.raises(RuntimeError, match="Exceed max numerical nugget"):
        la.chol_safe(M)
    la.set_num_nugget(0.0)
    try:
        with pytest.raises(RuntimeError, match="not strictly positive"):
            la.chol_safe(M)
    finally:
        la.set_num_nugget(1e-10)
    assert la.get_num_nugget() == 1e-10


def test_safe_chol_rcond_check_switch(la):
    # a well-conditioned block and a decoupled pivot 1e-20: potrf succeeds
    # exactly, rcond(L)^2 ~ 1e-20 < 1e-18 rejects the factor while the check
    # is on (one jitter step), and is accepted with the check off
    M = np.zeros((6, 6))
    M[:5, :5] = spd(5, 3, cond=10.0)
    M[5, 5] = 1e-20
    Lon, incon = la.chol_safe(M, return_inc=True)
    _, inco, _ = O.safe_chol_lower(M, O.default_cfg(rcond_check=True))
    assert incon == inco == 1
    la.check_chol_rcond(False)
    try:
        assert not la.chol_rcond_checked()
        _, incoff = la.chol_safe(M, return_inc=True)
        _, inco2, _ = O.safe_chol_lower(M, O.default_cfg(rcond_check=False))
        assert incoff == inco2 == 0
    finally:
        la.check_chol_rcond(True)
    print(f"rcond check on: inc {incon}; off: inc {incoff}")


def _dtrcon2(L):
    r, info = lapack.dtrcon(np.asfortranarray(L), norm="1", uplo="L", diag="N")
    assert info == 0
    return r * r


@pytest.mark.parametrize("n,cond", [(1, 1.0), (10, 1e2), (64, 1e6), (100, 1e10), (257, 1e4)])
def test_rcond_chol_matches_dtrcon(la, n, cond):
    L = np.linalg.cholesky(spd(n, 100 + n, cond=cond))
    r = la.rcond_chol(L)
    ref = _dtrcon2(L)
    print(f"rcond_chol n={n}: {r:.6e} dtrcon^2 {ref:.6e}")
    assert abs(r - ref) <= 1e-10 * ref


def test_rcond_chol_signs_and_singular(la):
    L = np.linalg.cholesky(spd(20, 8))
    D = np.where(np.arange(20) % 3 == 0, -1.0, 1.0)
    assert abs(la.rcond_chol(L * D) - la.rcond_chol(L)) <= 1e-14 * la.rcond_chol(L)
    assert abs(la.rcond_chol(L * D) - _dtrcon2(L * D)) <= 1e-10 * _dtrcon2(L * D)
    Z = L.copy()
    Z[7, 7] = 0.0
    assert la.rcond_chol(Z) == 0.0


def test_rcond_approx_chol(la):
    L = np.tril(np.random.default_rng(1).uniform(0.5, 2.0, (9, 9)))
    d = np.diag(L)
    assert la.rcond_approx_chol(L) == pytest.approx((d.min() / d.max()) ** 2, rel=1e-15)


def kernel_matrix(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, 3))
    return O.corr_mat("matern5_2", X, X, np.array([0.4, 0.5, 0.6]))


@pytest.mark.parametrize("n,no", [(2, 1), (70, 64), (70, 6), (200, 128), (300, 1), (300, 299), (257, 190)])
def test_chol_block_matches_oracle(la, n, no):
    C = kernel_matrix(n, n + no) if n > 10 else spd(n, 1)
    Loo = np.linalg.cholesky(C[:no, :no])
    L = la.chol_block(C, Loo)
    Lo, path = O.chol_block(C, Loo)
    assert path == "block"
    err = np.max(np.abs(L - Lo)) / np.max(np.abs(Lo))
    back = np.max(np.abs(L @ L.T - C))
    print(f"chol_block n={n} no={no}: rel err vs oracle {err:.2e}, |LL' - C| {back:.2e}")
    assert np.array_equal(L, np.tril(L))
    assert np.array_equal(L[:no, :no], np.tril(Loo))
    assert err < 1e-10
    assert back < 1e-12


def test_chol_block_falls_back(la):
    n, no = 90, 70
    C = spd(n, 5, cond=10.0)
    Loo = 0.1 * np.linalg.cholesky(C[:no, :no])
    assert O.chol_block(C, Loo)[1] == "full"
    L = la.chol_block(C, Loo)
    err = np.max(np.abs(L - np.linalg.cholesky(C)))
    print(f"chol_block fall-back: max err vs full factor {err:.2e}")
    assert err < 1e-12


def test_chol_block_full_old_block(la):
    C = spd(40, 2)
    Loo = np.linalg.cholesky(C)
    assert np.array_equal(la.chol_block(C, Loo), Loo)
