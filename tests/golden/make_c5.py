"""Fixture of the uneven NestedKriging configuration at full size
(N = 200 000 points in [0,1]^5, 256 groups of 300..1500 points, matern3_2,
constant trend; X, y and the groups from bench.nested_workload(), seed 127):
the CPU oracle's NK predictor (oracle/nested.py, NestedKriging.cpp:353-465)
at 8 of the workload's prediction points, with the prior fixed (theta,
sigma2, beta0 below: the fit's optimiser is not what this checks).  The
group-pair sums (N^2 / 2 kernel values) run over 8 processes.

Run: python tests/golden/make_c5.py   (a few minutes on 8 cores)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from oracle import nested as NKO  # noqa: E402

THETA = np.array([0.45, 0.5, 0.55, 0.6, 0.65])
SIGMA2 = 1.3
BETA0 = 0.4
POINTS = np.arange(8) * 1250  # rows of the 10 000 prediction points


def main():
    X, y, group_of, sizes, Xn = bench.nested_workload()
    groups = [np.flatnonzero(group_of == g) for g in range(len(sizes))]
    Xq = Xn[POINTS]
    mean, sd, info = NKO.predict_nk(X, y, groups, THETA, SIGMA2, BETA0, "matern3_2", Xq, return_info=True,
                                    workers=8)
    np.savez(os.path.join(HERE, "c5_nk_predict.npz"), theta=THETA, sigma2=SIGMA2, beta0=BETA0, points=POINTS,
             Xq=Xq, mean=mean, sd=sd, info=info)
    print("mean", mean, "sd", sd, "info", info)


if __name__ == "__main__":
    main()
