"""CPU study of the mixed-precision recSM factorization: which levels' FP32
products destroy the preconditioner at a given condition number.

The compressed operator (ACA factors, leaf blocks) comes from the oracle; the
factorization is restated here in numpy, level by level from the deepest, as
the HIP engine runs it (the node step of level d: Z = V^T U^SM over the
ancestors' and its own columns, the capacitance C = I + V^T U^SM from the own
columns, Y = C^-1 Z over the ancestors' columns, U^SM -= U^SM_own Y), with a
precision per level:

  "f64"  every operation in FP64;
  "f32"  the level's products in FP32 (operands rounded to FP32, numpy's
         float32 matmul accumulating in FP32, results rounded), its
         capacitance formed in FP64 from the FP32 products and factored in
         FP64 (as the engine's MIXED mode does);

and with ``store32`` the whole U^SM panel held in FP32 storage, as the
engine's MIXED mode stores it.  Every ACA term is rebalanced first (u s,
v / s, s = sqrt(|v| / |u|): the product is unchanged, and late terms of
partial-pivot ACA exceed the FP32 range otherwise).  Leaves are always factored and applied in FP64.  The
solve applies the (rounded) factors in FP64, which favours convergence: a
preconditioner that fails here fails on the device too.

Test infrastructure only (the oracle provides the compressed operator and
its FP64 product A_H x)."""
import numpy as np
import scipy.linalg as sla


class Hodlr:
    """The compressed operator of an oracle handle in arrays."""

    def __init__(self, o, points, family, ell, diag):
        from oracle import pyoracle
        self.o = o
        self.n = o.n
        o.compress()
        self.nodes = o.nodes()          # (level, lo, split, hi, kt, kb)
        self.leaves = o.leaves()        # (level, lo, hi)
        self.f = []
        for i in range(len(self.nodes)):
            kt, kb = int(self.nodes[i][4]), int(self.nodes[i][5])
            Ut, Vt = o.factor(i, 0)[:, :kt], o.factor(i, 1)[:, :kt]
            Ub, Vb = o.factor(i, 2)[:, :kb], o.factor(i, 3)[:, :kb]
            self.f.append((Ut, Vt, Ub, Vb))
        self.leaf_lu = []
        for lvl, lo, hi in self.leaves:
            m = int(hi - lo + 1)
            A = pyoracle.kernel_block(points, family, ell, diag, int(lo), int(lo), m, m)
            self.leaf_lu.append(sla.lu_factor(A))
        self.nlev = int(self.nodes[:, 0].max()) + 1 if len(self.nodes) else 0
        self.K = [0] * (self.nlev + 1)
        for lvl, lo, sp, hi, kt, kb in self.nodes:
            self.K[lvl] = max(self.K[lvl], int(kt), int(kb))
        self.off = np.concatenate([[0], np.cumsum(self.K)]).astype(int)

    def matvec(self, X):
        return self.o.matvec(X)


def _r32(a):
    return np.asarray(a, dtype=np.float32)


def _balance(U, V):
    nu = np.linalg.norm(U, axis=0)
    nv = np.linalg.norm(V, axis=0)
    s = np.where((nu > 0) & (nv > 0), np.sqrt(nv / np.where(nu > 0, nu, 1.0)), 1.0)
    return U * s, V / s


def factorize(H, prec, store32=False):
    """prec(level) -> "f64" | "f32".  Returns the factors for solve()."""
    N, off, K = H.n, H.off, H.K
    Ktot = int(off[H.nlev])
    W = np.zeros((N, max(Ktot, 1)))
    Vs = []
    for i, (lvl, lo, sp, hi, kt, kb) in enumerate(H.nodes):
        Ut, Vt, Ub, Vb = H.f[i]
        # (every level: an FP32 level rounds its ancestors' columns too)
        Ut, Vt = _balance(Ut, Vt)
        Ub, Vb = _balance(Ub, Vb)
        st = sp - lo + 1
        c0 = off[lvl]
        W[lo:lo + st, c0:c0 + kt] = Ut
        W[lo + st:hi + 1, c0:c0 + kb] = Ub
        Vs.append((Vt, Vb))
    if store32:
        W = _r32(W).astype(np.float64)

    def store(a):
        return _r32(a).astype(np.float64) if store32 else a

    C_lu = [None] * len(H.nodes)
    leaf_level = {}
    for j, (lvl, lo, hi) in enumerate(H.leaves):
        leaf_level.setdefault(int(lvl), []).append(j)
    deepest = max(int(H.leaves[:, 0].max()), H.nlev - 1)
    for d in range(deepest, -1, -1):
        for j in leaf_level.get(d, []):
            _, lo, hi = H.leaves[j]
            oc = int(off[min(d, H.nlev)])
            if oc > 0:
                W[lo:hi + 1, :oc] = store(sla.lu_solve(H.leaf_lu[j], W[lo:hi + 1, :oc]))
        for i in np.nonzero(H.nodes[:, 0] == d)[0]:
            lvl, lo, sp, hi, kt, kb = (int(v) for v in H.nodes[i])
            st = sp - lo + 1
            Vt, Vb = Vs[i]
            od = int(off[d])
            nc = od + K[d]
            top, bot = slice(lo, lo + st), slice(lo + st, hi + 1)
            f32 = prec(d) == "f32"
            if f32:
                Zt = (_r32(Vb).T @ _r32(W[top, :nc])).astype(np.float64)
                Zb = (_r32(Vt).T @ _r32(W[bot, :nc])).astype(np.float64)
            else:
                Zt = Vb.T @ W[top, :nc]
                Zb = Vt.T @ W[bot, :nc]
            Kn = kt + kb
            C = np.eye(Kn)
            C[:kt, kt:] = Zb[:, od:od + kb]
            C[kt:, :kt] = Zt[:, od:od + kt]
            C_lu[i] = sla.lu_factor(C)
            if od > 0:
                Y = sla.lu_solve(C_lu[i], np.vstack([Zb[:, :od], Zt[:, :od]]))
                if f32:
                    Y32 = _r32(Y)
                    W[top, :od] = (_r32(W[top, :od]) - _r32(W[top, od:od + kt]) @ Y32[:kt]).astype(np.float64)
                    W[bot, :od] = (_r32(W[bot, :od]) - _r32(W[bot, od:od + kb]) @ Y32[kt:]).astype(np.float64)
                else:
                    W[top, :od] = store(W[top, :od] - W[top, od:od + kt] @ Y[:kt])
                    W[bot, :od] = store(W[bot, :od] - W[bot, od:od + kb] @ Y[kt:])
    return dict(W=W, V=Vs, C=C_lu, deepest=deepest)


def solve(H, F, B):
    """x = M^-1 b with the factors in FP64: the leaves, then the node
    corrections level by level from the deepest."""
    X = np.array(B, dtype=np.float64, order="F").reshape(H.n, -1).copy()
    for j, (lvl, lo, hi) in enumerate(H.leaves):
        X[lo:hi + 1] = sla.lu_solve(H.leaf_lu[j], X[lo:hi + 1])
    W, off = F["W"], H.off
    for d in range(F["deepest"], -1, -1):
        for i in np.nonzero(H.nodes[:, 0] == d)[0]:
            lvl, lo, sp, hi, kt, kb = (int(v) for v in H.nodes[i])
            st = sp - lo + 1
            Vt, Vb = F["V"][i]
            top, bot = slice(lo, lo + st), slice(lo + st, hi + 1)
            od = int(off[d])
            rhs = np.vstack([Vt.T @ X[bot], Vb.T @ X[top]])
            Y = sla.lu_solve(F["C"][i], rhs)
            X[top] -= W[top, od:od + kt] @ Y[:kt]
            X[bot] -= W[bot, od:od + kb] @ Y[kt:]
    return X


def contraction(H, F, B):
    """max over columns of ||b - A_H M^-1 b|| / ||b||: the error reduction of
    one refinement step x += M^-1 r (below 1 it converges)."""
    X = solve(H, F, B)
    R = B - H.matvec(X)
    return float(np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)))


def fgmres(H, F, b, iters):
    """Right-preconditioned flexible GMRES from x = 0 (one cycle of `iters`),
    relative residual after each iteration."""
    n = H.n
    beta = np.linalg.norm(b)
    V = np.zeros((n, iters + 1))
    Z = np.zeros((n, iters))
    Hm = np.zeros((iters + 1, iters))
    V[:, 0] = b / beta
    out = []
    for j in range(iters):
        Z[:, j] = solve(H, F, V[:, j])[:, 0]
        w = H.matvec(Z[:, j:j + 1])[:, 0]
        for p in range(2):  # Gram-Schmidt, twice
            h = V[:, :j + 1].T @ w
            w -= V[:, :j + 1] @ h
            Hm[:j + 1, j] += h
        Hm[j + 1, j] = np.linalg.norm(w)
        if Hm[j + 1, j] > 0:
            V[:, j + 1] = w / Hm[j + 1, j]
        e1 = np.zeros(j + 2)
        e1[0] = beta
        y = np.linalg.lstsq(Hm[:j + 2, :j + 1], e1, rcond=None)[0]
        x = Z[:, :j + 1] @ y
        out.append(float(np.linalg.norm(b - H.matvec(x[:, None])[:, 0]) / beta))
    return out
