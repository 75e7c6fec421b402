"""Independent restatement of the reference's BFGS fit loop
(src/lib/Kriging.cpp:1904-2119) for checking GPU fits: the CPU oracle's LL
and gradient, scipy's fmin_l_bfgs_b (L-BFGS-B 3.0) as the optimiser, log
reparametrisation of theta, the restart rules and best-of-starts selection.

Test infrastructure only.
"""
import numpy as np
from scipy.optimize import fmin_l_bfgs_b

from oracle import oracle as O

EPS = np.finfo(float).eps


def mt19937_uniform(seed, count):
    """std::mt19937(seed) + std::uniform_real_distribution<double>{} as in
    libstdc++ (generate_canonical with two 32-bit draws): the reference's
    Random::randu stream (src/lib/Random.cpp:13-40)."""
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    idx = 624

    def nxt():
        nonlocal idx
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF

    out = []
    for _ in range(count):
        a = nxt()
        b = nxt()
        r = (a + b * 4294967296.0) / 18446744073709551616.0
        if r >= 1.0:
            r = np.nextafter(1.0, 0.0)
        out.append(r)
    return np.array(out)


def fit_one_start(P, th_start, lower, upper, max_iter=20, pgtol=1e-3, factr=1e10, max_restart=10, m=5):
    """One start of the BFGS loop with its restarts (Kriging.cpp:1950-2041):
    returns (objective -LL at the kept point, theta there)."""
    glo, gup = np.log(lower), np.log(upper)

    def fofn(g):
        th = np.exp(g)
        ll, gr, _ = P.log_likelihood(th)
        return -ll, -(gr * th)

    gamma = np.log(th_start)
    lo = np.minimum(gamma, glo)
    up = np.maximum(gamma, gup)
    best_f, best_g = np.inf, gamma.copy()
    retry = 0
    while retry <= max_restart:
        x, f, info = fmin_l_bfgs_b(fofn, gamma, bounds=list(zip(lo, up)), m=m, factr=factr, pgtol=pgtol,
                                   maxiter=max_iter)
        gamma = x
        if f < best_f:
            best_f, best_g = f, x.copy()
        sol_to_lb = np.min(np.abs(np.exp(x) - lower))
        task = info["task"] if isinstance(info["task"], str) else info["task"].decode()
        if retry < max_restart and (task.startswith("ABNORMAL_TERMINATION_IN_LNSRCH") or info["nit"] <= 2
                                    or sol_to_lb < EPS or f > best_f):
            rth = (th_start + lower) / 2.0 ** (retry + 1)
            gamma = np.log(rth)
            lo = np.minimum(gamma, lo)
            up = np.maximum(gamma, up)
            retry += 1
        else:
            break
    ll, _, _ = P.log_likelihood(np.exp(best_g), want_grad=False)
    return -ll, np.exp(best_g)


def fit_bfgs(X, y, kernel, theta0_rows, lower, upper, max_iter=20, pgtol=1e-3, factr=1e10, max_restart=10, m=5):
    """Multistart BFGS from the given starting rows, without nugget.
    Returns (best_theta, best_ll, per_start_objectives)."""
    P = O.OracleProblem(X, y, kernel, "constant")
    results = [fit_one_start(P, th, lower, upper, max_iter, pgtol, factr, max_restart, m)
               for th in np.atleast_2d(theta0_rows)]
    k = int(np.argmin([r[0] for r in results]))
    return results[k][1], -results[k][0], [r[0] for r in results]
