"""CPU tests pinning the rank-predictor oracle (oracle/mlp_oracle.py): its
gradient against central finite differences, GELU against closed forms,
bfloat16 rounding against hand-computed bit patterns, block features on a
configuration with known geometry."""
import math

import numpy as np

from oracle import mlp_oracle as M


def _params(seed=3, scale=0.3):
    return np.random.default_rng(seed).normal(0, scale, M.NPARAM)


def test_param_layout():
    assert M.NPARAM == 4801
    assert (M.OFF_B1, M.OFF_W2, M.OFF_B2, M.OFF_W3, M.OFF_B3) == (512, 576, 4672, 4736, 4800)


def test_gelu_closed_forms():
    x = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    # GELU(x) = x Phi(x); Phi(0) = 1/2, Phi(-x) = 1 - Phi(x)
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    assert np.allclose(M.gelu(x) - M.gelu(-x), x, atol=1e-15)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(x), fd, atol=1e-9)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    P = _params()
    X = rng.normal(0, 1, (40, 8))
    t = rng.normal(0, 1, 40)
    _, g = M.loss_and_grad(P, X, t)
    h = 1e-6
    for p in list(range(0, M.NPARAM, 97)) + [M.OFF_B1, M.OFF_W2 + 5, M.OFF_W3 + 3, M.OFF_B3]:
        e = np.zeros(M.NPARAM)
        e[p] = h
        fd = (M.loss_and_grad(P + e, X, t)[0] - M.loss_and_grad(P - e, X, t)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g[p])), (p, fd, g[p])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, 1.0 + 2**-9, -2.5, 3.14159265], dtype=np.float32)
    r = M.bf16(x)
    # 7 fraction bits: 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7 and rounds
    # to even (1.0); 1 + 3*2^-9 is above it (up); 1 + 2^-9 below it (down)
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2**-7, 1.0]
    assert r[4] == -2.5
    assert r[5] == np.float32(3.140625)


def test_block_features_known_geometry():
    pts = np.array([[0.0, 0.0], [0.2, 0.0], [1.0, 0.0], [1.2, 0.0]])
    f = M.block_features(pts, 0, 2, 2, 2, 3, 0.5, 1e-8, 1)
    assert abs(f[0] - math.log(4)) < 1e-15
    assert abs(f[1] - 1.0 / 0.2) < 1e-12   # centroids 0.1 and 1.1, radii 0.1 each
    assert abs(f[2] - 0.2) < 1e-12 and abs(f[3] - 0.2) < 1e-12
    assert abs(f[4] - math.log(0.5)) < 1e-15 and abs(f[5] + 8) < 1e-12
    assert f[6] == 3 and f[7] == 1
