"""Autotuner of the per-size-bucket launch table.

For each size bucket of the batched pipeline (largest matrix of at most 4,
8, 16, 32, 64, 128 64-row blocks, and the rest) it times batched LL +
gradient evaluations of a representative workload under every candidate
(Cholesky panel width P in block columns, stream parts S, trailing updates
on 64 x 64 or 128 x 64 tiles W) and keeps the fastest.  The table is written as JSON; a context loads it at creation from
tuning_<arch>.json beside liblkgpu.so (or the file LKGPU_TUNING names).

Run on the GPU the table is for:

    python -m libkriging_amd.autotune [--out PATH] [--reps R] [--quick]

Representative workloads: n = 64 max_nblk points (the bucket's largest
matrices), d = 6, matern5_2, constant trend, a batch of 64 slots up to
n = 1024, 16 at 2048, 4 at 4096, 1 above; parameter vectors log-uniform in
[0.1, 1].  Results do not depend on the table (each slot's arithmetic is the
same under every P and S); only the launch structure changes.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _native as N

BUCKETS = [4, 8, 16, 32, 64, 128, None]
PANELS = [1, 2, 4, 8]
STREAMS = [1, 2, 4]
WIDE = [0, 1]
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning_gfx950.json")


def _batch_size(n):
    return 64 if n <= 1024 else 16 if n <= 2048 else 4 if n <= 4096 else 1


def _table(panel, streams, wide):
    return {"buckets": [{"max_nblk": b, "panel_blocks": panel, "streams": streams, "wide": wide} for b in BUCKETS]}


def time_config(ctx, pid, theta, panel, streams, wide, reps):
    ctx.set_tuning(_table(panel, streams, wide))
    ids = [pid] * theta.shape[0]
    ctx.ll_grad_batch(ids, theta, True)  # capture the graph of this layout
    ctx.ll_grad_batch(ids, theta, True)
    ctx.synchronize()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.ll_grad_batch(ids, theta, True)
        ctx.synchronize()
        t.append(time.perf_counter() - t0)
    return float(np.median(t)) * 1e3


def tune(reps=5, quick=False, log=print):
    ctx = N.Context(0)
    ctx.set_graph(True)
    rng = np.random.default_rng(2024)
    d = 6
    out = []
    for b in BUCKETS:
        if b is None:
            # the unbounded bucket takes the choice of the 128-block one
            out.append(dict(out[-1], max_nblk=None))
            continue
        n = 64 * b
        if quick and n > 2048:
            out.append({"max_nblk": b, "panel_blocks": 4, "streams": 1 if _batch_size(n) < 8 else 4, "wide": 0,
                        "ms": None, "note": "not measured (--quick)"})
            continue
        B = _batch_size(n)
        X = rng.uniform(size=(n, d))
        y = np.sin(3 * X).sum(axis=1)
        pid = ctx.problem(X, y, np.ones((n, 1)), "matern5_2")
        theta = np.exp(rng.uniform(np.log(0.1), np.log(1.0), size=(B, d)))
        best = None
        results = {}
        for P in PANELS:
            if P > b:
                continue
            for S in STREAMS:
                if B % (8 * S) != 0 and S > 1:
                    continue  # this batch would not split into S parts
                for W in WIDE:
                    ms = time_config(ctx, pid, theta, P, S, W, reps)
                    results[f"P{P}_S{S}_W{W}"] = round(ms, 4)
                    if best is None or ms < best[0]:
                        best = (ms, P, S, W)
        ctx.drop_problem(pid)
        log(f"bucket <= {b} blocks (n = {n}, batch {B}): best P = {best[1]}, S = {best[2]}, W = {best[3]}, "
            f"{best[0]:.3f} ms; all {results}")
        out.append({"max_nblk": b, "panel_blocks": best[1], "streams": best[2], "wide": best[3], "ms": round(best[0], 4),
                    "n": n, "batch": B, "candidates_ms": results})
    ctx.close()
    return {"arch": "gfx950", "workload": "LL+grad, matern5_2, d = 6, constant trend", "buckets": out}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="only buckets up to n = 2048")
    args = ap.parse_args()
    table = tune(args.reps, args.quick)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
