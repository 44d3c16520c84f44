"""Writes tests/golden/matern2d_65536_oracle.npz: a compact summary of the
oracle's recSM on the N=65536 Matern-3/2 workload (points uniform in the unit
square in kd order, seed 12345; l = 0.2, +1e-3 on the diagonal, leaf 128,
ACA tolerance 1e-10, 16 right-hand sides ~ N(0,1)): ACA ranks of every node,
the log-determinant, per-column solution norms and 512 seeded solution
entries.  The oracle needs about ten minutes on one core for this size, so
the full solution is not kept.

    python tests/golden/make_matern2d_summary.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dense_hodlr_amd import inputs as I  # noqa: E402
from oracle import pyoracle  # noqa: E402


def summarize(x, logdet, nodes, n):
    idx = np.sort(I.SplitMix64(777).take(512) % np.uint64(n)).astype(np.int64)
    return dict(nodes=np.asarray(nodes, dtype=np.int32), logdet=np.float64(logdet),
                colnorm=np.linalg.norm(x, axis=0), idx=idx, xs=x[idx, :])


def main():
    cfg = I.config_inputs("matern2d_65536")
    o = pyoracle.OracleHODLR(cfg["n"], leaf=cfg["leaf"], points=cfg["points"], family="matern32",
                             length_scale=0.2, diag=1e-3)
    o.set_LRTolerance(cfg["tol"])
    x = o.recSM_Solve(cfg["rhs"])
    out = summarize(x, o.logAbsDeterminant(), o.nodes(), cfg["n"])
    np.savez(os.path.join(HERE, "matern2d_65536_oracle.npz"), **out)
    print("logdet", out["logdet"])


if __name__ == "__main__":
    main()
