"""GPU tests of the library's device allocation cache (hodlr_memory_*):
a second handle of the same shape is served from the blocks the first one
released, and its results are bit-identical to the first handle's (reused
blocks hold stale data, so this also checks that nothing relies on fresh
allocations being zero)."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import HODLR_Matrix, HodlrNotConverged, memory_cache, memory_stats, memory_trim
from dense_hodlr_amd import inputs as I

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _run(pts, b, precision):
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=128)
    g.set_LRTolerance(1e-10)
    g.set_precision(precision)
    try:
        x = g.recSM_Solve(b)
    except HodlrNotConverged as e:  # the best iterate: what matters here is reuse
        x = e.x
    ld = g.logAbsDeterminant()
    nodes = g.nodes()
    del g
    return x, ld, nodes


@pytest.mark.parametrize("precision", ["fp64", "fp64_refined", "mixed"])
def test_reused_blocks_give_identical_results(precision):
    _case(f"memory cache reuse {precision}")
    n = 8192
    pts = I.points_2d_kd(n, 31)
    b = I.rhs_normal(n, 3, 32)
    memory_trim()
    x0, l0, n0 = _run(pts, b, precision)
    s0 = memory_stats()
    assert s0["cached_bytes"] > 0, s0
    x1, l1, n1 = _run(pts, b, precision)
    s1 = memory_stats()
    print("after first:", s0, "\nafter second:", s1, flush=True)
    assert s1["hits"] > s0["hits"]
    assert s1["misses"] == s0["misses"], "a handle of the same shape allocated again"
    assert np.array_equal(n0, n1)
    assert np.array_equal(x0, x1)
    assert l0 == l1


def test_trim_and_cache_off():
    _case("memory trim / cache off")
    n = 4096
    pts = I.points_2d_kd(n, 33)
    b = I.rhs_normal(n, 1, 34)
    x0, _, _ = _run(pts, b, "fp64")
    assert memory_stats()["cached_bytes"] > 0
    memory_trim()
    s = memory_stats()
    assert s["cached_bytes"] == 0 and s["live_bytes"] == 0, s
    memory_cache(False)
    try:
        x1, _, _ = _run(pts, b, "fp64")
        assert memory_stats()["cached_bytes"] == 0
    finally:
        memory_cache(True)
    assert np.array_equal(x0, x1)
