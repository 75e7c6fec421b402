#!/usr/bin/env python3
"""Headline benchmark: batched FP64 LL + adjoint-gradient evaluations per
second on one MI355X, against the plain one-evaluation-per-launch GPU path.

Workload (the 64-start multistart fit configuration): n = 1024 design points
X ~ U[0,1)^6, y = prod_j (1 + 0.5 cos(3 x_j)) + N(0, 1e-4), exp kernel,
constant trend, 64 parameter vectors theta ~ log-U[0.05, 2]^6 -- all drawn
with numpy's PCG64 from a fixed seed.  One step = one batched evaluation of
the 64 LL + gradients (Cholesky, L^-1, GLS solves, R^-1 and the adjoint
gradient, all FP64) with X and y resident in HBM.

Prints ONE JSON line (rank 0, N = 1).  Extra fields: plain_evals_per_s (the
same 64 evaluations as 64 single-item launches) and batched_vs_plain.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_POINTS, DIM, NSTART = 1024, 6, 64
SEED = 126  # 123 + index of the multistart configuration
MFMA_FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20


def workload(seed=SEED, n=N_POINTS, d=DIM, nstart=NSTART):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.prod(1.0 + 0.5 * np.cos(3.0 * X), axis=1) + rng.normal(0.0, 1e-2, size=n)
    theta = np.exp(rng.uniform(np.log(0.05), np.log(2.0), size=(nstart, d)))
    return X, y, theta


def host_cpu():
    """(model name, cores this process may use) of the host"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(X, y, theta, gpu_ll, gpu_grad, budget_s=10.0, max_evals=20):
    """The reference algorithm on the host: the CPU oracle (a C restatement
    with the reference's serial pair loops) calling LAPACK's dpotrf / dtrtrs /
    dtrcon as the reference does through Armadillo (scipy's OpenBLAS build),
    timed on a bounded sample of the same 64 evaluations, first with one
    BLAS thread (the reference's multistart fit runs max(1, nproc / 64)
    threads, Kriging.cpp:1744-1746) and then with every core of this
    process; also the restated-loop oracle of the parity tests, and the
    parity of the GPU results on the sample."""
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    model, ncores = host_cpu()
    P = O.OracleProblem(X, y, "exp", "constant")

    def timed(threads, lapack, limit):
        O.use_lapack(lapack)
        ll_err, g_err = 0.0, 0.0
        k = 0
        with threadpool_limits(threads):
            t0 = time.perf_counter()
            while k < limit and (time.perf_counter() - t0) < budget_s:
                ll, g, _ = P.log_likelihood(theta[k])
                ll_err = max(ll_err, abs(gpu_ll[k] - ll) / abs(ll))
                g_err = max(g_err, np.linalg.norm(gpu_grad[k] - g) / np.linalg.norm(g))
                k += 1
            dt = time.perf_counter() - t0
        O.use_lapack(False)
        return k / dt, k, ll_err, g_err

    try:
        v1, k1, e1, g1 = timed(1, True, max_evals)
        va, ka, ea, ga = timed(ncores, True, max_evals)
        vr, kr, er, gr = timed(1, False, 5)
    finally:
        O.use_lapack(False)
    sample = (f"{k1} of the {theta.shape[0]} LL+grad evaluations (n={X.shape[0]}, d={X.shape[1]}): C oracle "
              f"with the reference's serial pair loops, LAPACK dpotrf/dtrtrs/dtrcon (scipy OpenBLAS), 1 BLAS thread")
    cb = {"value": v1, "unit": "evals/s", "cores": 1, "kind": "port", "sample": sample,
          "cpu_model": model, "host_cores": ncores,
          "all_cores": {"value": va, "cores": ncores, "evaluations": ka},
          "restated_loops_1_thread": {"value": vr, "evaluations": kr}}
    return cb, max(e1, ea, er), max(g1, ga, gr)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_batch64_n1024_d6.txt")


def lauum_traffic_bytes():
    """HBM bytes per k_lauum_adjoint launch of this workload, from the
    copy-calibrated FETCH_SIZE / WRITE_SIZE passes summarised in profiles/"""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("k_lauum_adjoint traffic_bytes_per_launch"):
                    return int(line.split()[-1])
    except OSError:
        pass
    return None


def mfma_probe_sweep():
    """FP64 MFMA issue rate from registers (the probe library) over
    independent chains per wave and workgroups (256 threads) per launch, with
    the mean shader clock of the workgroups"""
    from libkriging_amd import probe
    print(f"{'chains':>6s} {'workgroups':>10s} {'ms':>9s} {'TFLOP/s':>8s} {'clock MHz':>9s}")
    for chains in (1, 2, 4, 8):
        for blocks in (256, 512, 1024, 2048):
            ms, tf, mhz = probe.mfma(blocks, 20000 if blocks <= 1024 else 10000, chains)
            print(f"{chains:6d} {blocks:10d} {ms:9.3f} {tf:8.2f} {mhz:9.0f}", flush=True)


def cholesky_large(N, n=8192, d=10, reps=3):
    """The single-evaluation configuration at n = 8192 (gauss kernel, d = 10,
    theta log-U[0.3, 1.5], nugget ratio 1e-6): the Cholesky timed live over
    its own launches (n^3 / 3 flops per factorisation against the FP64 MFMA
    peak), the L^-1 stage likewise (n^3 / 3 flops), and the whole LL +
    gradient (11 components) by wall clock."""
    rng = np.random.default_rng(125)
    X = rng.uniform(size=(n, d))
    y = rng.normal(size=n)
    theta = np.exp(rng.uniform(np.log(0.3), np.log(1.5), size=d))
    gam = np.concatenate([theta, [1.0 / (1.0 + 1e-6)]])[None, :]
    ctx = N.Context(0)
    pid = ctx.problem(X, y, np.ones((n, 1)), "gauss", "nugget")
    ctx.ll_grad_batch([pid], gam, False)  # pool allocation and graph capture
    ctx.set_timed_stage("cholesky")
    for _ in range(reps):
        _, _, st, _ = ctx.ll_grad_batch([pid], gam, False)
    ms, count = ctx.take_stage_time()
    ctx.set_timed_stage("trtri")
    for _ in range(reps):
        ctx.ll_grad_batch([pid], gam, True)
    ms_tri, _ = ctx.take_stage_time()
    ctx.set_timed_stage(-1)
    assert (st == 0).all()
    ctx.ll_grad_batch([pid], gam, True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.ll_grad_batch([pid], gam, True)
    ctx.synchronize()
    ms_llg = (time.perf_counter() - t0) / reps * 1e3
    tflops = n ** 3 / 3.0 / (ms * 1e-3) / 1e12
    tri_tf = n ** 3 / 3.0 / (ms_tri * 1e-3) / 1e12
    ctx.drop_problem(pid)
    return {"n": n, "avg_ms": ms, "runs": count, "achieved_tflops": tflops,
            "frac_of_fp64_mfma_peak": tflops / MFMA_FP64_PEAK_TFLOPS,
            "trtri_ms": ms_tri, "trtri_frac_of_fp64_mfma_peak": tri_tf / MFMA_FP64_PEAK_TFLOPS,
            "ll_grad_ms": ms_llg}


def nested_workload(seed=127, n=200_000, d=5, groups=256):
    """The uneven NestedKriging configuration: X ~ U[0,1)^(n x d),
    y = sum_j sin(4 x_j) + x_1 x_5 + N(0, 1e-4); 256 group sizes drawn
    lognormal (sigma_log 0.5 around n / 256), clipped to [300, 1500] and
    adjusted by +-1 until they sum to n; points assigned by a seeded
    permutation."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(n, d))
    y = np.sum(np.sin(4.0 * X), axis=1) + X[:, 0] * X[:, d - 1] + rng.normal(0.0, 1e-2, size=n)
    sizes = np.clip(np.round(rng.lognormal(np.log(n / groups), 0.5, size=groups)), 300, 1500).astype(np.int64)
    while sizes.sum() != n:
        k = rng.integers(groups)
        step = 1 if sizes.sum() < n else -1
        if 300 <= sizes[k] + step <= 1500:
            sizes[k] += step
    perm = rng.permutation(n)
    group_of = np.empty(n, dtype=np.int32)
    o = 0
    for g, sz in enumerate(sizes):
        group_of[perm[o:o + sz]] = g
        o += sz
    Xn = rng.uniform(size=(10_000, d))
    return X, y, group_of, sizes, Xn


def bucket_fractions(X, y, group_of, theta, reps=5):
    """LL+grad of the NestedKriging groups at theta, batched by size bucket
    (64-row blocks per group: 5..8, 9..16, 17..24) and all 256 groups in one
    uneven batch: evaluations/s and the FP64 MFMA fraction of n_g^3 flops
    per evaluation (n^3 / 3 each for the Cholesky, L^-1 and the R^-1 tiles)."""
    from libkriging_amd import _native as N
    ctx = N.Context(0)
    ng = np.bincount(group_of)
    pids = []
    for g in range(len(ng)):
        idx = np.where(group_of == g)[0]
        pids.append(ctx.problem(X[idx], y[idx], np.ones((len(idx), 1)), "matern3_2"))
    nblk = (ng + 63) // 64
    sets = {"nblk 5-8": [g for g in range(len(ng)) if nblk[g] <= 8],
            "nblk 9-16": [g for g in range(len(ng)) if 9 <= nblk[g] <= 16],
            "nblk 17-32": [g for g in range(len(ng)) if 17 <= nblk[g] <= 32],
            "all groups": list(range(len(ng)))}
    out = {}
    for name, sel in sets.items():
        if not sel:
            continue
        ids = [pids[g] for g in sel]
        th = np.repeat(np.asarray(theta, dtype=np.float64)[None, :], len(sel), axis=0)
        for _ in range(2):
            ctx.ll_grad_batch(ids, th, True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.ll_grad_batch(ids, th, True)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / reps
        tf = float(np.sum(ng[sel].astype(np.float64) ** 3)) / dt / 1e12
        out[name] = {"groups": len(sel), "ms": 1e3 * dt, "evals_per_s": len(sel) / dt, "tflops": tf,
                     "frac_of_fp64_mfma_peak": tf / MFMA_FP64_PEAK_TFLOPS, "parts": ctx.last_parts()}
    for p in pids:
        ctx.drop_problem(p)
    ctx.close()
    return out


def nested_leg(points=2000):
    """The uneven NestedKriging configuration: fit (all 256 groups' BFGS in
    shared batched pipelines) and NK predict on `points` of the 10 000
    prediction points, the pair-contraction kernel timed live with HIP
    events: 2 sum_{i<j} n_i n_j flops per point."""
    import libkriging_amd as lk
    X, y, group_of, sizes, Xn = nested_workload()
    Xq = Xn[:points]
    nk = lk.NestedKriging("matern3_2")
    t0 = time.perf_counter()
    nk.fit_groups(y, X, group_of, "constant", "BFGS")
    t_fit = time.perf_counter() - t0
    fit_evals, fit_batches = nk.fit_statistics()
    nk.predict(Xq[:64], True)  # warm
    nk.time_pairs(True)
    t0 = time.perf_counter()
    mean, sd = nk.predict(Xq, True)
    t_pred = time.perf_counter() - t0
    pair_ms, pair_launches, pair_points = nk.pair_time()
    nk.time_pairs(False)
    buckets = bucket_fractions(X, y, group_of, nk.theta())
    s = sizes.astype(np.float64)
    pair_flops = (s.sum() ** 2 - np.sum(s ** 2)) * pair_points
    achieved = pair_flops / (pair_ms * 1e-3) / 1e12
    return {
        "config": {"n": int(s.sum()), "d": 5, "groups": len(sizes), "group_min": int(sizes.min()),
                   "group_max": int(sizes.max()), "points": points, "kernel": "matern3_2"},
        "fit_s": t_fit, "fit_evaluations": int(fit_evals), "fit_batches": int(fit_batches),
        "fit_evals_per_s": fit_evals / t_fit,
        "nk_points_per_s": points / t_pred, "predict_s": t_pred, "predict_fallbacks": nk.predict_fallbacks(),
        "nk_pair": {"bound": "mfma", "achieved": achieved, "peak": MFMA_FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / MFMA_FP64_PEAK_TFLOPS, "kernel": "k_nk_pair", "ms": pair_ms,
                    "launches": pair_launches},
        "ll_grad_by_bucket": buckets,
        "theta": nk.theta().tolist(), "sigma2": nk.sigma2(), "beta0": nk.beta0(),
        "finite": bool(np.all(np.isfinite(mean)) and np.all(np.isfinite(sd))),
    }


def run_nested(args):
    """The nested configuration on its own line: value = NK points/s."""
    r = nested_leg(args.nested_points)
    line = {
        "metric": "NestedKriging NK predict points/s (N=200000, 256 uneven groups, d=5, matern3_2)",
        "value": r["nk_points_per_s"], "unit": "points/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * r["predict_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": dict(r["config"], workload="NestedKriging uneven groups: fit (BFGS per group) + NK predict"),
        "roofline": r["nk_pair"],
    }
    line.update({k: v for k, v in r.items() if k not in ("config", "nk_pair")})
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chol", action="store_true", help="skip the n = 8192 Cholesky measurement")
    ap.add_argument("--no-nested", action="store_true", help="skip the NestedKriging fit + NK predict leg")
    ap.add_argument("--workload", default="multistart", choices=["multistart", "nested"],
                    help="multistart: the headline batched LL+grad; nested: NestedKriging fit + NK predict")
    ap.add_argument("--nested-points", type=int, default=2000, help="prediction points of the nested workload")
    ap.add_argument("--stage", default="adjoint", help="kernel timed live for the roofline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help=f"also run a {COPY_BYTES >> 20} MiB device copy (counter calibration)")
    ap.add_argument("--mfma-probe-sweep", action="store_true",
                    help="only print the FP64 MFMA issue-rate probe over chains x workgroups")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the multistart batch would shard by starts, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.mfma_probe_sweep:
        mfma_probe_sweep()
        return
    if args.workload == "nested":
        run_nested(args)
        return

    from libkriging_amd import _native as N

    X, y, theta = workload()
    n, d = X.shape
    ctx = N.Context(0)
    pid = ctx.problem(X, y, np.ones((n, 1)), "exp")
    ids = [pid] * NSTART

    # batched path: one hipGraph replay per step, the batch split into the
    # parts (one stream each) the launch table gives this size bucket
    for _ in range(args.warmup):
        ctx.ll_grad_batch(ids, theta, True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ll, grad, st, inc = ctx.ll_grad_batch(ids, theta, True)
    ctx.synchronize()
    t_batched = time.perf_counter() - t0
    parts = ctx.last_parts()
    assert (st == 0).all(), "an evaluation failed"
    evals_s = NSTART * args.steps / t_batched

    # roofline kernel timed live with HIP events around its launch, in a loop
    # of its own: a timed run keeps the whole batch on one stream (with the
    # batch split, a part's launch shares the GPU with the other parts' and its
    # duration says nothing about the kernel)
    roof_steps = max(5, min(args.steps, 20))
    ctx.ll_grad_batch(ids, theta, True)
    ctx.set_timed_stage(args.stage)
    for _ in range(roof_steps):
        ll_t, grad_t, _, _ = ctx.ll_grad_batch(ids, theta, True)
    stage_ms, stage_count = ctx.take_stage_time()
    ctx.set_timed_stage(-1)
    split_equals_single = bool(np.array_equal(ll_t, ll) and np.array_equal(grad_t, grad))

    # plain path: the same evaluations, one per launch sequence
    for b in range(2):
        ctx.ll_grad_batch([pid], theta[b:b + 1], True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        for b in range(NSTART):
            l1, g1, _, _ = ctx.ll_grad_batch([pid], theta[b:b + 1], True)
    ctx.synchronize()
    t_plain = time.perf_counter() - t0
    plain_evals_s = NSTART * args.plain_steps / t_plain if args.plain_steps > 0 else float("nan")
    # same numbers from both paths
    l1, g1, _, _ = ctx.ll_grad_batch([pid], theta[NSTART - 1:NSTART], True)
    same = bool(l1[0] == ll[NSTART - 1] and np.array_equal(g1[0], grad[NSTART - 1]))

    # roofline of the live-timed kernel
    if args.stage == "adjoint":
        # R^-1 = L^-T L^-1 on MFMA: n^3/3 flops per evaluation
        flops = NSTART * n ** 3 / 3.0
        achieved = flops / (stage_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": MFMA_FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / MFMA_FP64_PEAK_TFLOPS, "traffic": lauum_traffic_bytes(),
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated (profiles/)",
                "kernel": "k_lauum_adjoint", "avg_ms": stage_ms, "launches": stage_count}
    elif args.stage == "build":
        nbytes = NSTART * (8.0 * n * (n + 1) / 2 + 8.0 * n * d)
        achieved = nbytes / (stage_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_build",
                "avg_ms": stage_ms, "launches": stage_count}
    else:
        roof = {"bound": "mfma", "achieved": None, "peak": MFMA_FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": None, "traffic": None, "stage": args.stage, "avg_ms": stage_ms}

    roof["timed_batches"] = roof_steps
    line = {
        "metric": "LL+grad evals/s (batched adjoint vs plain GPU), n=1024 d=6, rel err<=1e-8",
        "value": evals_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_batched / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "64-start batched LL+grad, n=1024 d=6 exp kernel, constant trend",
                   "n": n, "d": d, "kernel": "exp", "batch": NSTART, "seed": SEED},
        "plain_evals_per_s": plain_evals_s,
        "batched_vs_plain": evals_s / plain_evals_s,
        "plain_equals_batched": same,
        "streams": parts,
        "split_equals_single_stream": split_equals_single,
        "roofline": roof,
    }
    from libkriging_amd import probe
    copy_ms = probe.device_copy(COPY_BYTES, 5)
    line["hbm_copy_GBps"] = 2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9  # read + write of a plain copy kernel
    # FP64 MFMA rate reachable from registers alone (8 independent
    # accumulation chains per wave, 8 waves per SIMD): the practical ceiling
    # beside the published 78.6 TFLOP/s, and the shader clock it ran at
    _, tf, mhz = probe.mfma(2048, 10000, 8)
    line["mfma_f64_probe_TFLOPs"] = tf
    line["mfma_f64_probe_clock_MHz"] = mhz
    if not args.no_chol:
        line["cholesky_n8192"] = cholesky_large(N)
    if not args.no_nested:
        line["nested_c5"] = nested_leg(args.nested_points)
    if not args.no_cpu:
        cb, ll_err, g_err = cpu_baseline(X, y, theta, ll, grad)
        line["cpu_baseline"] = cb
        line["parity_vs_oracle"] = {"ll_rel_max": ll_err, "grad_rel_max": g_err}
    print(json.dumps(line))
    if args.copy_calibration:
        # a plain device copy of known size, run in the same process so that
        # a counter-collection pass over this command also measures it:
        # FETCH_SIZE / WRITE_SIZE of this kernel calibrate those counters
        from libkriging_amd import probe
        ms = probe.device_copy(COPY_BYTES, 3)  # k_copy: reads and writes COPY_BYTES per launch
        print(json.dumps({"copy_calibration_bytes": COPY_BYTES, "copy_ms": ms,
                          "copy_GBps": 2 * COPY_BYTES / (ms * 1e-3) / 1e9}), file=sys.stderr)


if __name__ == "__main__":
    main()
