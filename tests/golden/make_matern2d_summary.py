"""Writes tests/golden/matern2d_65536_oracle.npz: a compact summary of the
oracle's recSM on the N=65536 Matern-3/2 workload (points uniform in the unit
square in kd order, seed 12345; l = 0.2, +1e-3 on the diagonal, leaf 128,
ACA tolerance 1e-10, 16 right-hand sides ~ N(0,1)): ACA ranks of every node,
the log-determinant, per-column solution norms and 512 seeded solution
entries, of the plain recSM solution and of the solution refined by three
corrections with double-double residuals (which converges to the solution of
A_H x = b rounded, independent of the order of the solver's sums).  The
oracle needs about fifteen minutes on one core for this size, so the full
solutions are not kept.

    python tests/golden/make_matern2d_summary.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dense_hodlr_amd import inputs as I  # noqa: E402
from oracle import pyoracle  # noqa: E402


def summarize(x, xr, steps, res, logdet, nodes, n):
    idx = np.sort(I.SplitMix64(777).take(512) % np.uint64(n)).astype(np.int64)
    return dict(nodes=np.asarray(nodes, dtype=np.int32), logdet=np.float64(logdet),
                colnorm=np.linalg.norm(x, axis=0), idx=idx, xs=x[idx, :],
                colnorm_refined=np.linalg.norm(xr, axis=0), xs_refined=xr[idx, :],
                refine_steps=np.int32(steps), refine_residual=np.float64(res))


def main():
    cfg = I.config_inputs("matern2d_65536")
    o = pyoracle.OracleHODLR(cfg["n"], leaf=cfg["leaf"], points=cfg["points"], family="matern32",
                             length_scale=0.2, diag=1e-3)
    o.set_LRTolerance(cfg["tol"])
    x = o.recSM_Solve(cfg["rhs"])
    xr, steps, res = o.refined_solve(cfg["rhs"], target=0.0, max_steps=3)
    out = summarize(x, xr, steps, res, o.logAbsDeterminant(), o.nodes(), cfg["n"])
    np.savez(os.path.join(HERE, "matern2d_65536_oracle.npz"), **out)
    print("logdet", out["logdet"], "refined residual", res, "after", steps, "corrections",
          "refined vs plain", np.linalg.norm(xr - x) / np.linalg.norm(xr))


if __name__ == "__main__":
    main()
