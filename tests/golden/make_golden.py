"""Writes the golden fixtures of tests/golden/ with numpy and scipy.

These are textbook routines, not the reference: scipy.linalg.lu_factor
(LAPACK getrf: partial pivoting, first maximum) pins the oracle's LU and
numpy.linalg.svd pins its singular values.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20240607)
    lu = {}
    for n in (1, 2, 5, 16, 33, 64):
        A = rng.uniform(-1, 1, (n, n))
        F, piv = sla.lu_factor(A)
        lu[f"A{n}"] = A
        lu[f"LU{n}"] = F
        lu[f"piv{n}"] = piv.astype(np.int32)
    # a matrix with ties in the pivot column: first maximum must win
    A = np.array([[1.0, 2.0, 3.0], [-4.0, 1.0, 0.5], [4.0, -2.0, 1.0]])
    F, piv = sla.lu_factor(A)
    lu["Atie"], lu["LUtie"], lu["pivtie"] = A, F, piv.astype(np.int32)
    np.savez(os.path.join(HERE, "lu_golden.npz"), **lu)

    sv = {}
    for (m, n) in ((8, 5), (5, 8), (20, 20), (40, 12)):
        A = rng.standard_normal((m, n))
        sv[f"A{m}x{n}"] = A
        sv[f"s{m}x{n}"] = np.linalg.svd(A, compute_uv=False)
    # prescribed spectrum for the absolute-threshold rank rule
    Q1, _ = np.linalg.qr(rng.standard_normal((12, 6)))
    Q2, _ = np.linalg.qr(rng.standard_normal((9, 6)))
    s = np.array([10.0, 1.0, 1e-3, 1e-9, 1e-12, 0.0])
    sv["Aspec"] = Q1 @ np.diag(s) @ Q2.T
    sv["sspec"] = np.linalg.svd(sv["Aspec"], compute_uv=False)
    np.savez(os.path.join(HERE, "svd_golden.npz"), **sv)
    print("wrote lu_golden.npz, svd_golden.npz")


if __name__ == "__main__":
    main()
