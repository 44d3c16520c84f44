"""Benchmark of the HODLR hot path on one MI355X.

One step = compression (ACA, fused kernel assembly) + recSM factorization +
solve of the 2D Gaussian workload: N = 262144 points uniform in the unit
square in kd order, K(x,y) = exp(-|x-y|^2 / 0.05^2) + 1e-4 on the diagonal,
leaf 256, ACA tolerance 1e-10, 64 right-hand sides ~ N(0,1).  Inputs (points
and right-hand sides) are resident in device memory before the timed region.

Prints one JSON line (the contract documented in the repository), with the
roofline of the dominant kernel (ACA) and a CPU baseline measured on a
bounded sample by the oracle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor+solve time at N=256K to 1e-10 rel. residual: mixed+learned vs plain FP64"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=262144)
    p.add_argument("--nrhs", type=int, default=64)
    p.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-n", type=int, default=8192)
    return p.parse_args()


def workload(n, nrhs, seed=12345):
    from dense_hodlr_amd import inputs as I
    return I.points_2d_kd(n, seed), I.rhs_normal(n, nrhs, seed + 1)


def gpu_step(pts, dB, dX, n, nrhs, precision):
    import ctypes as C

    from dense_hodlr_amd import HODLR_Matrix, _lib

    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    g.set_LRTolerance(1e-10)
    if precision == "mixed":
        g.set_precision("mixed")
    t0 = time.perf_counter()
    _lib.check(_lib.lib().hodlr_solve_device(g._h, _lib.FACTOR_SM, C.c_void_p(dB), n, nrhs,
                                             C.c_void_p(dX), n), g._h)
    wall = time.perf_counter() - t0
    return g, wall


def cpu_baseline(n_sample, nrhs):
    from oracle import pyoracle
    pts, B = workload(n_sample, nrhs, seed=777)
    o = pyoracle.OracleHODLR(n_sample, leaf=256, points=pts, family="gaussian",
                             length_scale=0.05, diag=1e-4)
    o.set_LRTolerance(1e-10)
    t0 = time.perf_counter()
    o.recSM_Solve(B)
    t = time.perf_counter() - t0
    return t, o.times()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch

    n, nrhs = a.n, a.nrhs
    pts, B = workload(n, nrhs)
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")  # column-major n x nrhs
    dX = torch.empty_like(dB)
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        gpu_step(pts, dB.data_ptr(), dX.data_ptr(), n, nrhs, a.precision)
    torch.cuda.synchronize()
    times, st = [], None
    t_all = time.perf_counter()
    for _ in range(a.steps):
        g, wall = gpu_step(pts, dB.data_ptr(), dX.data_ptr(), n, nrhs, a.precision)
        st = g.stats()
        times.append(st["compress_time"] + st["factor_time"] + st["solve_time"])
        del g
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_all
    X = dX.cpu().numpy().T
    ms = 1e3 * elapsed / a.steps
    value = float(np.median(times))

    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"gauss2d_kd N={n} l=0.05 diag=1e-4 leaf=256 tol=1e-10 nrhs={nrhs}",
                   "precision": a.precision},
        "phases_s": {"compress": st["compress_time"], "factor": st["factor_time"],
                     "solve": st["solve_time"]},
        "ranks_max_per_level": st["max_rank"],
        "aca_iterations": st["aca_iterations"],
    }
    if not a.no_cpu_baseline:
        t, parts = cpu_baseline(a.cpu_n, nrhs)
        out["cpu_baseline"] = {"value": t, "unit": "s", "cores": 1, "kind": "port",
                               "sample": f"oracle recSM compress+factor+solve, same kernel/leaf/tol, "
                                         f"N={a.cpu_n}, {nrhs} RHS, single thread",
                               "phases_s": parts}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
