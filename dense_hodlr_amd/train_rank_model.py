"""Regenerates the shipped rank-predictor weights (rank_mlp.npz) on one GPU.

Training data: seeded two-dimensional kernel systems of the large-system
kind (points uniform in the unit square in kd order, leaf 256), Gaussian or
Matern-3/2 with length scale log-uniform in [0.02, 0.3], N log-uniform in
[16384, 262144], ACA tolerance log-uniform in [1e-12, 1e-6], diagonal shift
1e-3 or 1e-4.  Every run is
compressed by the plain adaptive ACA; each off-diagonal block contributes
its 8 features and log(rank).  The network (8 -> 64 -> 64 -> 1, GELU) is
trained by Adam in the library's HIP kernels on standardized features (the
standardization is folded into the first layer afterwards); runs whose index
is a multiple of 10 are held out for validation.

    python -m dense_hodlr_amd.train_rank_model --runs 2000 --epochs 60

prints one JSON line (sample counts, losses, held-out rank errors) and
writes the weights.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import inputs as I
from .rank_model import WEIGHTS, RankModel, collect_samples


def training_runs(count: int, seed: int = 2024, first: int = 0):
    g = I.SplitMix64(seed)
    for r in range(count):
        u = g.uniform(5)
        if r < first:
            continue
        n = int(np.exp(np.log(16384) + (np.log(262144) - np.log(16384)) * u[0]))
        n = int(min(262144, max(16384, n)))
        family = "gaussian" if u[1] < 0.5 else "matern32"
        ell = float(np.exp(np.log(0.02) + (np.log(0.3) - np.log(0.02)) * u[2]))
        tol = float(10.0 ** (-6.0 - 6.0 * u[3]))
        # the shift only enters the leaves' diagonal, not the off-diagonal
        # blocks whose ranks are learned; both values of the workloads occur
        diag = 1e-4 if u[4] < 0.5 else 1e-3
        yield dict(run=r, n=n, points=I.points_2d_kd(n, seed * 1000 + r), family=family,
                   length_scale=ell, diag=diag, tol=tol, leaf=256)


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--runs", type=int, default=2000)
    p.add_argument("--epochs", type=int, default=60)
    p.add_argument("--batch", type=int, default=1024)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--out", default=WEIGHTS)
    p.add_argument("--first", type=int, default=0,
                   help="collect runs first .. runs-1 only (the earlier ones come from --samples-in)")
    p.add_argument("--samples-out", default=None,
                   help="save the collected samples (npz) and exit without training")
    p.add_argument("--samples-in", nargs="*", default=[],
                   help="samples saved by earlier --samples-out calls, added before training")
    a = p.parse_args()
    t0 = time.perf_counter()
    parts = {False: ([], []), True: ([], [])}
    for path in a.samples_in:
        with np.load(path, allow_pickle=False) as f:
            for held in (False, True):
                parts[held][0].append(f["X_val" if held else "X_train"])
                parts[held][1].append(f["t_val" if held else "t_train"])
    for r in training_runs(a.runs, a.seed, a.first):
        X, t = collect_samples([r])
        xs, ts = parts[r["run"] % 10 == 0]
        xs.append(X)
        ts.append(t)
        print(f"[run {r['run']}] n={r['n']} {r['family']} l={r['length_scale']:.3f} "
              f"tol={r['tol']:.1e} diag={r['diag']:g}: {len(t)} blocks, max rank "
              f"{int(np.exp(t.max()) + 0.5)}", flush=True)
    if a.samples_out:
        np.savez(a.samples_out,
                 X_train=np.concatenate(parts[False][0]), t_train=np.concatenate(parts[False][1]),
                 X_val=np.concatenate(parts[True][0]), t_val=np.concatenate(parts[True][1]))
        print(json.dumps({"samples_out": a.samples_out, "collect_s": time.perf_counter() - t0}))
        return
    Xt, tt = np.concatenate(parts[False][0]), np.concatenate(parts[False][1])
    Xv, tv = np.concatenate(parts[True][0]), np.concatenate(parts[True][1])
    t_collect = time.perf_counter() - t0
    # train on standardized features, then fold the standardization into the
    # first layer (W1 / sd, b1 - (W1 / sd) mu) so the network takes raw features
    mu = Xt.mean(0)
    sd = Xt.std(0)
    sd[sd == 0] = 1.0
    m = RankModel(seed=a.seed)
    t1 = time.perf_counter()
    losses = m.train((Xt - mu) / sd, tt, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    t_train = time.perf_counter() - t1
    P = m.params.astype(np.float64)
    W1 = P[:512].reshape(64, 8) / sd
    P[512:576] -= W1 @ mu
    P[:512] = W1.ravel()
    m.params = P.astype(np.float32)
    yv = m.predict_bf16(Xv)
    err = np.abs(yv - tv)
    m.save(a.out)
    print(json.dumps({
        "train_samples": int(len(tt)), "val_samples": int(len(tv)),
        "collect_s": t_collect, "train_s": t_train,
        "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
        "val_mean_abs_log_rank_error": float(err.mean()),
        "val_frac_within_25pct": float(np.mean(np.exp(err) <= 1.25)),
        "val_frac_under_predicted_25pct_margin": float(np.mean(np.exp(tv) > 1.25 * np.exp(yv))),
        "weights": a.out}), flush=True)


if __name__ == "__main__":
    main()
