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
FP64_PEAK_TFS = 78.6
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
                        "(e.g. 4,0,4096,32,1,4)")
    p.add_argument("--cpu-n", type=int, nargs="+", default=[2048, 4096, 8192, 16384],
                   help="oracle sample sizes (at least 2; the largest is the reported sample)")
    p.add_argument("--no-iterative", action="store_true",
                   help="skip the GMRES solve against the exact kernel")
    p.add_argument("--profile-factor", action="store_true",
                   help="only compress the workload, then run one FP64 recSM factorization "
                        "with its kernel groups timed, print them and exit (for counter runs: "
                        "every GEMM, LU and solve kernel of the process is the factorization's)")
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
        g.set_launch_shapes(*SHAPES[:4], bool(SHAPES[4]), SHAPES[5] if len(SHAPES) > 5 else 3)
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


def work_counts(nodes, n, nrhs):
    """Algorithmic work of the recSM path on a tree with these nodes
    (level, min_i, split, max_i, rank_top, rank_bottom), ranges not among
    them being leaves:
      aca_mac  sum over blocks of (m+n) k(k+1)/2  (residual multiply-adds),
      aca_ent  sum over blocks of (m+n)(k+1)      (kernel entries evaluated),
      factor   flops of the leaf LUs, the leaf and node solves of the
               ancestors' U^SM columns, the V^T U^SM products, the node
               matrices' LUs and the U^SM updates,
      solve    flops of one solve with nrhs columns."""
    rng = {(int(r[1]), int(r[3])): (int(r[2]), int(r[4]), int(r[5])) for r in nodes}
    w = dict(aca_mac=0.0, aca_ent=0.0, factor=0.0, solve=0.0)
    stack = [(0, n - 1, 0)]
    while stack:
        lo, hi, anc = stack.pop()
        if (lo, hi) not in rng:  # a leaf
            s_ = hi - lo + 1
            w["factor"] += 2.0 / 3.0 * s_ ** 3 + 2.0 * s_ * s_ * anc
            w["solve"] += 2.0 * s_ * s_ * nrhs
            continue
        sp, kt, kb = rng[(lo, hi)]
        st, sb, K = sp - lo + 1, hi - sp, kt + kb
        for m, nn, k in ((st, sb, kt), (sb, st, kb)):
            w["aca_mac"] += (m + nn) * k * (k + 1) / 2.0
            w["aca_ent"] += (m + nn) * (k + 1.0)
        prod = kb * st + kt * sb      # the V^T U^SM products' rows x ranks
        upd = kt * st + kb * sb       # the updates' rows x ranks
        w["factor"] += (2.0 * prod * (anc + K) + 2.0 / 3.0 * K ** 3 + 2.0 * K * K * anc
                        + 2.0 * upd * anc)
        w["solve"] += nrhs * (2.0 * prod + 2.0 * K * K + 2.0 * upd)
        stack.append((lo, sp, anc + kt))
        stack.append((sp + 1, hi, anc + kb))
    return w


def cpu_baseline(sizes, nrhs):
    """The oracle's recSM (compress + factor + solve, one thread) at the
    sample sizes of the same system, with each run's phase times and work
    counts; the largest is the reported sample."""
    from oracle import pyoracle
    res = []
    for n in sizes:
        pts, B = workload(n, nrhs, seed=777)
        o = pyoracle.OracleHODLR(n, leaf=256, points=pts, family="gaussian",
                                 length_scale=0.05, diag=1e-4)
        o.set_LRTolerance(1e-10)
        t0 = time.perf_counter()
        o.recSM_Solve(B)
        wall = time.perf_counter() - t0
        res.append(dict(n=n, wall=wall, phases=o.times(), work=work_counts(o.nodes(), n, nrhs)))
        del o
    return res


def fit_cpu(runs):
    """Least-squares fit of the oracle's phase times to its work counts:
    compress = a aca_mac + b aca_ent (non-negative), factor = c factor,
    solve = d solve.  Returns the coefficients and a predictor."""
    from scipy.optimize import nnls
    Ac = np.array([[r["work"]["aca_mac"], r["work"]["aca_ent"]] for r in runs])
    tc = np.array([r["phases"]["compress"] for r in runs])
    sc = Ac.max(axis=0)
    ab, _ = nnls(Ac / sc, tc)
    ab = ab / sc
    coef = {"compress_per_mac": ab[0], "compress_per_entry": ab[1]}
    for ph, key in (("sm_factor", "factor"), ("sm_solve", "solve")):
        x = np.array([r["work"][key] for r in runs])
        t = np.array([r["phases"][ph] for r in runs])
        coef[key + "_per_flop"] = float(x @ t / (x @ x))

    def predict(w):
        return (coef["compress_per_mac"] * w["aca_mac"] + coef["compress_per_entry"] * w["aca_ent"]
                + coef["factor_per_flop"] * w["factor"] + coef["solve_per_flop"] * w["solve"])
    return {k: float(v) for k, v in coef.items()}, predict


def main():
    global SHAPES
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
    if a.profile_factor:
        SHAPES = [int(v) for v in a.launch_shapes.split(",")] if a.launch_shapes else None
        pts, _ = workload(a.n, 1)
        g = make_handle(pts, a.target, "fp64", 0)
        g.storeLRinTree()
        g.set_profiling(True)
        g.recSM_Factorize()
        st = g.stats()
        print(json.dumps({"workload": f"gauss2d_kd N={a.n} recSM factorization",
                          "factor_s": st["factor_time"], "groups": st["factor_groups"],
                          "max_rank": st["max_rank"]}), flush=True)
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
    nodes_wl = g.nodes()
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
        "timed_path": "FP64-preconditioned refinement to 1e-10 with learned panel sizing (the "
                      "metric's mixed leg does not converge here: see mixed; plain_fp64 is the "
                      "same refined path without the predictor; fp64_unrefined is the "
                      "reference's plain recSM_Solve, which stops near 1e-6)",
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
                      "mean_abs_log_rank_error": st["pred_log_err"],
                      "panel_growths": st["aca_panel_grows"],
                      "training": "2000 seeded runs (rank_mlp.npz, train_rank_model.py)"},
        "step_walls_s": walls,
        "plain_fp64": {"label": "refined, no predictor",
                       "value": plain, "step_walls_s": pwalls, "compress_s": pst["compress_time"],
                       "factor_s": pst["factor_time"], "solve_and_refine_s": pst["solve_time"],
                       "residual": pst["last_residual"], "panel_growths": pst["aca_panel_grows"]},
        "speedup_vs_plain": plain / value,
        "tuning": tuning,
    }
    from dense_hodlr_amd import memory_stats
    out["device_memory_cache"] = memory_stats()

    # the reference's own path: plain FP64 recSM_Solve, no refinement
    uwalls = []
    for _ in range(a.steps):
        g = make_handle(pts, a.target, "fp64", 0)
        uwalls.append(solve_device(g, dB.data_ptr(), dX.data_ptr(), n, nrhs))
        su = g.stats()
        Xu = dX.cpu().numpy().T.copy(order="F")
        ru = float(np.max(np.linalg.norm(g.residual(B, Xu), axis=0) / bn))
        g = None
    out["fp64_unrefined"] = {"label": "plain recSM_Solve (no refinement, no predictor)",
                             "value": float(np.median(uwalls)), "step_walls_s": uwalls,
                             "compress_s": su["compress_time"], "factor_s": su["factor_time"],
                             "solve_s": su["solve_time"], "residual": ru}

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
        times, grows, pbytes = {}, {}, {}
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
            grows[learned] = sd["aca_panel_grows"]
            pbytes[learned] = sd["aca_panel_bytes"]
            g = None
        out["learned_deploy"] = {
            "workload": f"matern32 2D kd N={n} l=0.1 leaf=256 tol=1e-10 (compression)",
            "plain_aca_s": min(times[False]), "predictor_sized_s": min(times[True]),
            "speedup": min(times[False]) / min(times[True]),
            "blocks": sd["pred_blocks"], "misses": sd["pred_misses"],
            "miss_rate": sd["pred_misses"] / max(1, sd["pred_blocks"]),
            "mean_abs_log_rank_error": sd["pred_log_err"], "max_rank": max(sd["max_rank"]),
            "panel_growths": {"plain": grows[False], "predictor": grows[True]},
            "panel_bytes": {"plain": pbytes[False], "predictor": pbytes[True]}}

    if not a.no_roofline:
        # the sustained ceilings of this device (probe kernels), beside the
        # datasheet peaks the fractions below are priced against
        from dense_hodlr_amd import device_peaks
        peaks = device_peaks()
        out["device_peaks_measured"] = dict(
            peaks, note="probe kernels on this device (hodlr_measure_device_peaks): streaming "
                        "read of 1 GiB, 512 MiB copy, back-to-back v_mfma_f64_16x16x4, v_fma_f64 "
                        "and v_mfma_f32_16x16x4 from registers; best of five launches")
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
            "frac_of_measured_read": achieved / peaks["hbm_read_gbs"],
        }
        # the factorization of the same handle, its kernel groups timed with
        # HIP events on the streams they run on (FP64: 78.6 TF/s on the
        # vector units and on the matrix cores alike)
        g.recSM_Factorize()
        fg = g.stats()["factor_groups"]
        names = {"gemm": "gemm_mfma_kernel<double> (grouped: V^T U^SM products and U^SM updates of "
                         "every node of a level)",
                 "lu": "batched LU (lu_panel_reg_kernel + lu_step_kernel + trailing GEMM), "
                       "leaves and node matrices",
                 "leaf_solve": "leaf LU solves of the ancestors' U^SM columns (the tuned "
                               "variant: lu_solve_mfma_kernel on the matrix cores by default)",
                 "node_solve": "node-matrix LU solves (lu_solve_mfma_kernel up to 256 rows, "
                               "lu_solve_wave/blk kernels above)"}
        more = []
        for key, label in names.items():
            q = fg[key]
            if q["time"] <= 0:
                continue
            tf = q["flops"] / q["time"] / 1e12
            more.append({"kernel": label, "bound": "mfma" if key == "gemm" else "fp64",
                         "achieved": tf, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                         "frac": tf / FP64_PEAK_TFS, "time_s": q["time"], "flops": q["flops"],
                         "bytes": q["bytes"], "launches": q["launches"],
                         "roofline_time_s": q["bound_time"],
                         "frac_of_roofline_time": q["bound_time"] / q["time"],
                         # (the batched LU runs on the vector units, the rest on the
                         # matrix cores with the default launch shapes)
                         "frac_of_measured_fp64": tf / (peaks["fp64_valu_tflops"] if key == "lu"
                                                        else peaks["fp64_mfma_tflops"])})
        out["roofline_factorization"] = more
        del g

    if not a.no_iterative:
        # GMRES against the exact kernel (evaluated on the device), left
        # preconditioned by the FP64 recSM solve at the same ACA tolerance:
        # the solution's residual against K itself, which the direct and
        # refined solves above only reach to the compression's distance
        from dense_hodlr_amd import HODLR_Matrix, HodlrNotConverged, _lib
        g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
        if SHAPES:
            g.set_launch_shapes(*SHAPES[:4], bool(SHAPES[4]), SHAPES[5] if len(SHAPES) > 5 else 3)
        t0 = time.perf_counter()
        status = "converged"
        try:
            _lib.check(_lib.lib().hodlr_iterative_solve_device(
                g._h, C.c_void_p(dB.data_ptr()), n, nrhs, C.c_void_p(dX.data_ptr()), n, 50,
                1e-10, 1e-10, _lib.LR_METHODS["partialPiv_ACA"], _lib.FACTOR_SM), g._h)
        except HodlrNotConverged:
            status = "HODLR_ERR_NOT_CONVERGED"
        wall = time.perf_counter() - t0
        si = g.stats()
        Xi = dX.cpu().numpy().T.copy(order="F")
        out["iterative_exact"] = {
            "label": "iterative_Solve: GMRES(32) on the exact kernel, recSM preconditioner at "
                     "ACA tolerance 1e-10, stop at preconditioned residual 1e-10",
            "status": status, "wall_s": wall, "iterations": si["iter_iterations"],
            "preconditioned_residual": si["iter_precond_residual"],
            "residual_vs_kernel": si["iter_residual"],
            "kernel_spot_check": kernel_spot_check(pts, B, Xi)}
        g = None

    gpu_sample = None
    if not a.no_cpu_baseline:
        # the GPU on the CPU sample itself: the same plain recSM_Solve
        ns = max(a.cpu_n)
        spts, sB = workload(ns, nrhs, seed=777)
        sdB = torch.from_numpy(np.ascontiguousarray(sB.T)).to("cuda")
        sdX = torch.empty_like(sdB)
        sw = []
        for _ in range(3):
            from dense_hodlr_amd import HODLR_Matrix
            g = HODLR_Matrix.from_kernel(spts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
            g.set_LRTolerance(1e-10)
            torch.cuda.synchronize()
            sw.append(solve_device(g, sdB.data_ptr(), sdX.data_ptr(), ns, nrhs))
            g = None
        gpu_sample = {"n": ns, "value_s": min(sw), "path": "plain recSM_Solve on the GPU"}
    if not a.no_cpu_baseline:
        runs = cpu_baseline(sorted(a.cpu_n), nrhs)
        last = runs[-1]
        coef, predict = fit_cpu(runs)
        fit_err = max(abs(predict(r["work"]) - r["wall"]) / r["wall"] for r in runs)
        hold = None
        if len(runs) >= 3:
            _, pred_h = fit_cpu(runs[:-1])
            hold = {"fit_sizes": [r["n"] for r in runs[:-1]], "predicted_n": last["n"],
                    "predicted_s": pred_h(last["work"]), "measured_s": last["wall"]}
        wl = work_counts(nodes_wl, n, nrhs)
        out["cpu_baseline"] = {
            "value": last["wall"], "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle recSM compress+factor+solve (no refinement), same kernel, leaf, "
                      f"tolerance and {nrhs} RHS at N={last['n']}, one thread",
            "phases_s": last["phases"],
            "sizes_s": {str(r["n"]): r["wall"] for r in runs},
            "gpu_same_sample": gpu_sample,
            "workload_estimate_s": predict(wl),
            "workload_estimate_kind": "fit, not measured: the oracle's phase times fitted to "
                                      "its algorithmic work counts at the sample sizes, "
                                      "evaluated on the workload's own ranks",
            "fit": {"model": "compress = a sum (m+n)k(k+1)/2 + b sum (m+n)(k+1); factor = c "
                             "flops; solve = d flops", "coefficients": coef,
                    "max_rel_residual": fit_err, "holdout": hold},
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
