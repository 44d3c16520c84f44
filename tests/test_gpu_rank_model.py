"""GPU tests of the learned rank predictor: the HIP forward, backward
(gradients) and BF16 inference against the numpy oracle, the block-feature
kernel against the oracle's features, training lowering the loss, and the
predictor leaving ACA's results untouched while sizing its panels."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import HODLR_Matrix, RankModel
from dense_hodlr_amd import inputs as I
from dense_hodlr_amd import rank_model as R
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 8)).astype(np.float32)
    t = (0.5 * X[:, 0] - 0.3 * X[:, 1] ** 2 + 0.1 * X[:, 2] * X[:, 3]).astype(np.float32)
    return X, t


def test_forward_and_gradient_match_oracle():
    _case("mlp forward/gradient vs oracle")
    m = RankModel(seed=11)
    P = m.params.astype(np.float64)
    X, t = _data(3000, 1)  # spans several gradient chunks of 512 samples
    y = m.forward(X)
    yo = M.forward(P, X.astype(np.float64))
    assert np.max(np.abs(y - yo)) <= 1e-5 * max(1.0, np.max(np.abs(yo)))
    loss, g = m.gradient(X, t)
    lo, go = M.loss_and_grad(P, X.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * lo
    err = np.abs(g - go)
    assert np.max(err) <= 1e-4 * np.max(np.abs(go)), (np.argmax(err), np.max(err))
    # deterministic: the same gradient bit for bit on a second call
    assert np.array_equal(m.gradient(X, t)[1], g)


def test_bf16_inference_matches_oracle():
    _case("mlp bf16 inference vs oracle")
    m = RankModel(seed=5)
    X, _ = _data(1000, 2)
    yb = m.predict_bf16(X)
    yo = M.forward_bf16(m.params, X)
    y32 = M.forward(m.params.astype(np.float64), X.astype(np.float64))
    # FP32 vs FP64 accumulation of the same bfloat16 operands; one rounding
    # flip of an activation moves y by at most ~2^-8 |W3| |a2|
    assert np.max(np.abs(yb - yo)) <= 2e-2 * max(1.0, np.max(np.abs(yo)))
    assert np.mean(np.abs(yb - yo)) <= 2e-3
    # bfloat16 is a genuine approximation of the FP32 network
    assert 0 < np.mean(np.abs(yb - y32)) <= 5e-2


def test_training_lowers_loss():
    _case("mlp training")
    m = RankModel(seed=3)
    X, t = _data(4096, 4)
    losses = m.train(X, t, epochs=40, batch=256, lr=3e-3, seed=9)
    assert losses[-1] < 0.2 * losses[0], losses[[0, -1]]
    # the trained parameters reproduce the reported loss
    lo, _ = M.loss_and_grad(m.params.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    assert abs(lo - losses[-1]) <= 1e-3 * max(lo, 1e-6) + 1e-6


def test_block_features_match_oracle():
    _case("block features")
    n = 8192
    pts = I.points_2d_kd(n, 3)
    g = HODLR_Matrix.from_kernel(pts, "matern32", 0.1, 1e-3, sizeThreshold=256)
    F = R.block_features(g, 1e-8)
    nd = g.nodes()
    assert F.shape == (2 * len(nd), 8)
    for i in (0, 1, 5, len(nd) - 1):
        lvl, lo, sp, hi = (int(v) for v in nd[i, :4])
        top = M.block_features(pts, lo, sp - lo + 1, sp + 1, hi - sp, lvl, 0.1, 1e-8, 1)
        bot = M.block_features(pts, sp + 1, hi - sp, lo, sp - lo + 1, lvl, 0.1, 1e-8, 1)
        assert np.allclose(F[2 * i], top, rtol=1e-5, atol=1e-6), (i, F[2 * i], top)
        assert np.allclose(F[2 * i + 1], bot, rtol=1e-5, atol=1e-6)


def test_predictor_sizes_panels_without_changing_results():
    _case("predictor attached vs plain ACA")
    n = 16384
    pts = I.points_2d_kd(n, 8)
    b = I.rhs_normal(n, 2, 9)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    g.set_LRTolerance(1e-10)
    x0 = g.recSM_Solve(b)
    ranks0 = g.nodes()
    m = RankModel.load() if __import__("os").path.exists(R.WEIGHTS) else RankModel(seed=1)
    g.set_rank_predictor(m, 0.25)
    x1 = g.recSM_Solve(b)
    st = g.stats()
    assert st["pred_blocks"] == 2 * len(ranks0)
    assert 0 <= st["pred_misses"] <= st["pred_blocks"]
    # ACA decides the ranks and the factors: identical results
    assert np.array_equal(g.nodes(), ranks0)
    assert np.array_equal(x1, x0)
    g.set_rank_predictor(None)
    assert np.array_equal(g.recSM_Solve(b), x0)
