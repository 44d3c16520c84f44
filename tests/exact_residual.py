"""Exact residual B - A_H X of a HODLR operator in rational arithmetic
(fractions.Fraction), from the FP64 values the operator is made of: dense
leaf blocks and the U, V factors of every node (read through an object with
nodes() and factor(idx, which), the GPU handle or the oracle), rounded once.
Used by the residual tests of the GPU library and of the oracle."""
from fractions import Fraction

import numpy as np


def _leaves(nodes, n):
    """Leaf intervals: children of non-leaf nodes that are not nodes."""
    if len(nodes) == 0:
        return [(0, n - 1)]
    ranges = {(int(lo), int(hi)) for _, lo, _, hi, _, _ in nodes}
    out = []
    for _, lo, sp, hi, _, _ in nodes:
        for a, b in ((int(lo), int(sp)), (int(sp) + 1, int(hi))):
            if (a, b) not in ranges:
                out.append((a, b))
    return sorted(out)


def exact_residual(g, A, B, X):
    n, nrhs = B.shape
    nodes = g.nodes()
    Y = [[Fraction(0)] * nrhs for _ in range(n)]
    Xf = [[Fraction(float(X[i, j])) for j in range(nrhs)] for i in range(n)]
    for a, b in _leaves(nodes, n):
        for i in range(a, b + 1):
            row = [Fraction(float(v)) for v in A[i, a:b + 1]]
            for j in range(nrhs):
                Y[i][j] += sum(r * Xf[a + q][j] for q, r in enumerate(row))
    for idx, (_, lo, sp, hi, kt, kb) in enumerate(nodes):
        lo, sp, hi = int(lo), int(sp), int(hi)
        for (rows0, cols0, which_u, which_v, k) in ((lo, sp + 1, 0, 1, kt), (sp + 1, lo, 2, 3, kb)):
            if k == 0:
                continue
            U = g.factor(idx, which_u)
            V = g.factor(idx, which_v)
            T = [[sum(Fraction(float(V[r, c])) * Xf[cols0 + r][j] for r in range(V.shape[0]))
                  for j in range(nrhs)] for c in range(k)]
            for r in range(U.shape[0]):
                for j in range(nrhs):
                    Y[rows0 + r][j] += sum(Fraction(float(U[r, c])) * T[c][j] for c in range(k))
    return np.array([[float(Fraction(float(B[i, j])) - Y[i][j]) for j in range(nrhs)]
                     for i in range(n)])
