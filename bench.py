"""Benchmark of the HODLR hot path on one MI355X.

Workload (the N = 262144 two-dimensional Gaussian system): points uniform in
the unit square in kd order, K(x,y) = exp(-|x-y|^2 / 0.05^2) + 1e-4 on the
diagonal, leaf 256, ACA tolerance 1e-10, 64 right-hand sides ~ N(0,1).
Points and right-hand sides are resident in device memory before the timed
region.

One step = one hodlr_solve_device call on a fresh handle: compression (ACA
with fused kernel assembly; U/V panels sized by the learned rank predictor),
recSM factorization, solve, and flexible-GMRES refinement against the FP64
HODLR operator until every column reaches a relative residual of 1e-10
(evaluated in double-double arithmetic).  The preconditioner is the FP64
recSM factorization: the plain recSM solution alone stops near 1e-6 on this
system (condition number ~1e7).  Beside it: the same solve without the
predictor (the plain FP64 path), the mixed-precision factorization (FP32
node products) run once with its outcome, and the predictor deployed on a
Matern-3/2 system it was not trained on.

Prints one JSON line: the contract fields, the roofline of the dominant
kernels (the ACA row and column residual passes, HBM bound, timed with HIP
events on the engine's stream), and a CPU baseline measured by the oracle on
a bounded sample.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor+solve time at N=256K to 1e-10 rel. residual: mixed+learned vs plain FP64"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "aca_gauss2d_262144_traffic.txt")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=262144)
    p.add_argument("--nrhs", type=int, default=64)
    p.add_argument("--target", type=float, default=1e-10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-mixed", action="store_true")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--no-learned-deploy", action="store_true")
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--launch-shapes", default=None,
                   help="fix the launch shapes instead of autotuning: ACA rows per thread, "
                        "GEMM tile rows, split-K target, LU panel, MFMA, leaf-solve variant "
                        "(e.g. 4,0,4096,32,1,0)")
    p.add_argument("--cpu-n", type=int, nargs=2, default=[4096, 8192])
    p.add_argument("--workload", default="gauss2d_262144", choices=["gauss2d_262144", "gp_batch_512"],
                   help="gp_batch_512: the batched Gaussian-process systems instead of the "
                        "headline system (a separate JSON line, not the headline metric)")
    p.add_argument("--copy-kernel", action="store_true",
                   help="only run a device-to-device copy of 1 GiB five times (calibrates the "
                        "HBM byte counters of rocprofv3 --pmc) and exit")
    return p.parse_args()


def copy_kernel():
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8,
                      "launches": 5}), flush=True)


from dense_hodlr_amd import inputs as I  # noqa: E402


def workload(n, nrhs, seed=12345):
    return I.points_2d_kd(n, seed), I.rhs_normal(n, nrhs, seed + 1)


SHAPES = None  # fixed launch shapes (--launch-shapes), applied to every handle


def kernel_spot_check(pts, B, X, rows=2048, seed=777, length=0.05, shift=1e-4, device="cuda"):
    """Residual of the solution against the exact kernel matrix on `rows`
    seeded rows S (no HODLR approximation): max over columns of
    ||B_S - K[S,:] X|| / ||B_S||, with the rows of K (Gaussian exp(-r^2/l^2)
    plus the diagonal shift) evaluated on the fly in FP64 on the GPU.  It
    measures the compressed operator's distance to the kernel matrix (ACA
    tolerance per block) as seen by this solution, not the solver."""
    import torch
    n = pts.shape[0]
    S = np.sort(np.argsort(I.SplitMix64(seed).uniform(n), kind="stable")[:rows])
    P = torch.from_numpy(np.ascontiguousarray(pts)).to(device)
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(device)
    R = np.empty((rows, X.shape[1]))
    for c0 in range(0, rows, 256):
        idx = torch.from_numpy(S[c0:c0 + 256]).to(device)
        d2 = ((P[idx, None, :] - P[None, :, :]) ** 2).sum(-1)
        K = torch.exp(-d2 / (length * length))
        K[torch.arange(len(idx), device=device), idx] += shift
        R[c0:c0 + 256] = (K @ Xd).cpu().numpy()
        del d2, K
    BS = B[S]
    return float(np.max(np.linalg.norm(BS - R, axis=0) / np.linalg.norm(BS, axis=0)))


def make_handle(pts, target, precision, max_steps):
    from dense_hodlr_amd import HODLR_Matrix
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    if SHAPES:
        g.set_launch_shapes(*SHAPES[:4], bool(SHAPES[4]), SHAPES[5] if len(SHAPES) > 5 else 0)
    g.set_LRTolerance(1e-10)
    g.set_precision(precision)
    g.set_refinement(target, max_steps)
    return g


def solve_device(g, dB, dX, n, nrhs):
    from dense_hodlr_amd import _lib
    t0 = time.perf_counter()
    _lib.check(_lib.lib().hodlr_solve_device(g._h, _lib.FACTOR_SM, C.c_void_p(dB), n, nrhs,
                                             C.c_void_p(dX), n), g._h)
    return time.perf_counter() - t0


def read_traffic():
    """HBM bytes per ACA iteration (row + column launch) from the committed
    rocprofv3 --pmc summary, or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    return vals.get("hbm_bytes_per_iteration")


def gp_batch(a):
    """512 independent GP covariance systems (sizes log-uniform in [1024,
    16384], 1D sorted points, squared exponential with l ~ U[0.01, 0.2], noise
    1e-2, leaf 64, ACA tol 1e-10), one right-hand side and the log-determinant
    each, every instance refined to 1e-10: one batched handle, per step."""
    import torch

    from dense_hodlr_amd import HODLR_Batch
    cfg = I.config_inputs("gp_batch_512")
    B = np.concatenate(cfg["rhs"])
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")
    dX = torch.empty_like(dB)
    out = {"workload": "gp_batch_512: 512 SE systems, N log-uniform [1024,16384], "
                       f"total rows {int(sum(cfg['n']))}, leaf 64, tol 1e-10, 1 RHS + logdet each"}
    for prec in ("fp64_refined", "mixed"):
        walls, ld_walls, st, res = [], [], None, None
        for it in range(a.warmup + a.steps):
            g = HODLR_Batch(cfg["points"], "se", cfg["length_scale"], cfg["diag"],
                            sizeThreshold=cfg["leaf"])
            g.set_LRTolerance(cfg["tol"])
            g.set_precision(prec)
            g.set_refinement(1e-10, 40)
            g.set_graphs(True)
            torch.cuda.synchronize()
            w = solve_device(g, dB.data_ptr(), dX.data_ptr(), len(B), 1)
            t0 = time.perf_counter()
            lds = g.logAbsDeterminants()
            tl = time.perf_counter() - t0
            if it >= a.warmup:
                walls.append(w)
                ld_walls.append(tl)
                st = g.stats()
                X = dX.cpu().numpy().T.copy(order="F")
                R = g.residual(B, X)
                res = max(np.linalg.norm(r) / np.linalg.norm(b)
                          for r, b in zip(g.split(R), g.split(B)))
            g = None
        out[prec] = {"solve_s": float(np.median(walls)), "logdet_s": float(np.median(ld_walls)),
                     "compress_s": st["compress_time"], "factor_s": st["factor_time"],
                     "refine_steps": st["refine_iterations"],
                     "max_instance_residual": float(res)}
    print(json.dumps(out), flush=True)


def cpu_baseline(sizes, nrhs):
    """The oracle's recSM (compress + factor + solve, one thread) at two
    sample sizes of the same system; the larger is the reported sample, the
    pair gives the exponent of a power-law extrapolation to the workload."""
    from oracle import pyoracle
    res = []
    for n in sizes:
        pts, B = workload(n, nrhs, seed=777)
        o = pyoracle.OracleHODLR(n, leaf=256, points=pts, family="gaussian",
                                 length_scale=0.05, diag=1e-4)
        o.set_LRTolerance(1e-10)
        t0 = time.perf_counter()
        o.recSM_Solve(B)
        res.append((n, time.perf_counter() - t0, o.times()))
    return res


def main():
    a = parse()
    if a.copy_kernel:
        copy_kernel()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "one GPU only: this build runs the whole system on a single "
                                   "MI355X; --gpus must be 1"}))
        sys.exit(2)
    if a.workload == "gp_batch_512":
        gp_batch(a)
        return
    import torch

    n, nrhs = a.n, a.nrhs
    pts, B = workload(n, nrhs)
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")  # column-major n x nrhs
    dX = torch.empty_like(dB)
    torch.cuda.synchronize()

    from dense_hodlr_amd import RankModel
    model = RankModel.load()

    def step(precision="fp64_refined", max_steps=30, learned=True):
        g = make_handle(pts, a.target, precision, max_steps)
        if learned:
            g.set_rank_predictor(model, 0.25)
        wall = solve_device(g, dB.data_ptr(), dX.data_ptr(), n, nrhs)
        return g, wall

    # untimed: the autotuner picks launch shapes for this problem shape; every
    # later handle of the shape (both paths below) uses them
    tuning = None
    global SHAPES
    shapes = [int(v) for v in a.launch_shapes.split(",")] if a.launch_shapes else None
    SHAPES = shapes
    if shapes is not None:
        tuning = {"fixed": shapes}
    if shapes is None and not a.no_autotune:
        g = make_handle(pts, a.target, "fp64_refined", 30)
        g.autotune()
        tuning = g.stats()["tuning"]
        g = None
    for _ in range(a.warmup):
        g, _ = step()
        del g
    torch.cuda.synchronize()
    walls, st, g = [], None, None
    t_all = time.perf_counter()
    for _ in range(a.steps):
        g = None
        g, wall = step()
        st = g.stats()
        walls.append(wall)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_all
    ms = 1e3 * elapsed / a.steps
    value = float(np.median(walls))

    # independent check of the last solution: its residual against A_H, in
    # double-double arithmetic and, for comparison, in plain FP64
    X = dX.cpu().numpy().T.copy(order="F")
    bn = np.linalg.norm(B, axis=0)
    res = float(np.max(np.linalg.norm(g.residual(B, X), axis=0) / bn))
    res64 = float(np.max(np.linalg.norm(B - g.matvec(X), axis=0) / bn))
    g = None
    spot = kernel_spot_check(pts, B, X)

    # the plain FP64 path: the same solve without the rank predictor (panels
    # grown adaptively from 32 columns), same number of steps
    pwalls, pst = [], None
    for _ in range(a.steps):
        g, wall = step(learned=False)
        pst = g.stats()
        pwalls.append(wall)
        g = None
    plain = float(np.median(pwalls))

    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"gauss2d_kd N={n} gaussian l=0.05 shift=1e-4 leaf=256 "
                               f"aca_tol=1e-10 nrhs={nrhs} target={a.target:g}",
                   "path": "learned rank sizing + FP64 recSM + refinement x += M^-1 r with "
                           "double-double residuals (flexible GMRES when a step gains less "
                           "than 8x; the FP32 factorization does not converge at this "
                           "condition number: see mixed)"},
        "phases_s": {"compress": st["compress_time"], "factor": st["factor_time"],
                     "solve_and_refine": st["solve_time"]},
        "refine_steps": st["refine_iterations"],
        "residual_reported": st["last_residual"],
        "residual_checked": res,
        "residual_fp64_evaluation": res64,
        "kernel_spot_check": {
            "rows": 2048, "seed": 777, "relative_residual": spot,
            "meaning": "||b_S - K[S,:] x|| / ||b_S|| against the exact kernel on 2048 seeded rows "
                       "(rows of K evaluated on the fly); the targets above are against A_H"},
        "ranks_max_per_level": st["max_rank"],
        "aca_iterations": st["aca_iterations"],
        "predictor": {"blocks": st["pred_blocks"], "misses": st["pred_misses"],
                      "mean_abs_log_rank_error": st["pred_log_err"]},
        "step_walls_s": walls,
        "plain_fp64": {"value": plain, "step_walls_s": pwalls, "compress_s": pst["compress_time"],
                       "factor_s": pst["factor_time"], "solve_and_refine_s": pst["solve_time"],
                       "residual": pst["last_residual"]},
        "speedup_vs_plain": plain / value,
        "tuning": tuning,
    }
    from dense_hodlr_amd import memory_stats
    out["device_memory_cache"] = memory_stats()

    if not a.no_mixed:
        from dense_hodlr_amd import HodlrError, HodlrNotConverged
        g = make_handle(pts, a.target, "mixed", 20)
        t0 = time.perf_counter()
        try:
            solve_device(g, dB.data_ptr(), dX.data_ptr(), n, nrhs)
            status = "converged"
        except HodlrNotConverged:
            status = "HODLR_ERR_NOT_CONVERGED"
        except HodlrError as e:
            status = str(e)
        wall = time.perf_counter() - t0
        sm = g.stats()
        ok = status == "converged"
        out["mixed"] = {"reached_target": ok, "status": status, "wall_s": wall,
                        "residual": sm["last_residual"], "refine_steps": sm["refine_iterations"],
                        "factor_s": sm["factor_time"],
                        "speedup_vs_plain": (plain / wall) if ok else None}
        del g

    if not a.no_learned_deploy:
        # the predictor on a system it was not trained on: Matern-3/2, l = 0.1,
        # N = 262144, tolerance 1e-10; compression only, with and without it
        dpts = I.points_2d_kd(n, 4242)
        times = {}
        for learned in (False, True, False, True):
            from dense_hodlr_amd import HODLR_Matrix
            g = HODLR_Matrix.from_kernel(dpts, "matern32", 0.1, 1e-3, sizeThreshold=256)
            g.set_LRTolerance(1e-10)
            if learned:
                g.set_rank_predictor(model, 0.25)
            t0 = time.perf_counter()
            g.storeLRinTree()
            times.setdefault(learned, []).append(time.perf_counter() - t0)
            sd = g.stats()
            g = None
        out["learned_deploy"] = {
            "workload": f"matern32 2D kd N={n} l=0.1 leaf=256 tol=1e-10 (compression)",
            "plain_aca_s": min(times[False]), "predictor_sized_s": min(times[True]),
            "speedup": min(times[False]) / min(times[True]),
            "blocks": sd["pred_blocks"], "misses": sd["pred_misses"],
            "miss_rate": sd["pred_misses"] / max(1, sd["pred_blocks"]),
            "mean_abs_log_rank_error": sd["pred_log_err"], "max_rank": max(sd["max_rank"])}

    if not a.no_roofline:
        # one extra compression with HIP events around every ACA row/column
        # launch (the timed steps above run without them)
        g = make_handle(pts, a.target, "fp64_refined", 30)
        g.set_profiling(True)
        g.storeLRinTree()
        sp = g.stats()
        t_aca = sp["aca_row_time"] + sp["aca_col_time"]
        iters = max(1, sp["aca_launches"])
        achieved = sp["aca_bytes"] / t_aca / 1e9
        traffic = read_traffic()
        out["roofline"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "aca_row_kernel + aca_col_kernel (one ACA iteration over all active blocks)",
            "algorithmic_bytes_per_launch": sp["aca_bytes"] / iters,
            "avg_launch_us": {"row": 1e6 * sp["aca_row_time"] / iters,
                              "col": 1e6 * sp["aca_col_time"] / iters},
            "launches": iters,
        }
        del g

    if not a.no_cpu_baseline:
        pts_ = cpu_baseline(a.cpu_n, nrhs)
        (n0, t0, _), (n1, t1, parts) = pts_[0], pts_[-1]
        p = math.log(t1 / t0) / math.log(n1 / n0) if t1 > 0 and t0 > 0 and n1 != n0 else 2.0
        out["cpu_baseline"] = {
            "value": t1, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle recSM compress+factor+solve (no refinement), same kernel, leaf, "
                      f"tolerance and {nrhs} RHS at N={n1}, one thread",
            "phases_s": parts,
            "sizes_s": {str(k): v for k, v, _ in pts_},
            "extrapolated_to_workload_s": t1 * (n / n1) ** p,
            "extrapolation_exponent": p,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
