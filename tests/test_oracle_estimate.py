"""Pins the oracle's dataset estimates (min/max, histogram, ranks, moments)."""
import numpy as np
import pytest

import oracle as O


def test_min_max_brute_force():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((1000, 5)).astype(np.float32)
    lim = O.min_max(a)
    assert np.array_equal(lim[:, 0], a.min(0)) and np.array_equal(lim[:, 1], a.max(0))


def test_histogram_counts_by_definition():
    rng = np.random.default_rng(1)
    a = rng.uniform(-3, 5, (4096, 3)).astype(np.float32)
    lim = O.min_max(a)
    counts = O.histogram_counts(lim, a, 256).reshape(3, 256)
    for d in range(3):
        lo, hi = np.float32(lim[d, 0]), np.float32(lim[d, 1])
        # estimate.cl:17-18 by definition, float arithmetic
        b = ((a[:, d] - lo) / (hi - lo) * np.float32(256)).astype(np.float32)
        b = np.minimum(np.floor(b).astype(np.int64), 255)
        assert np.array_equal(np.bincount(b, minlength=256), counts[d])
    assert counts.sum(1).tolist() == [4096] * 3


def test_histogram_edges_saturate():
    lim = np.array([[0.0, 1.0]], np.float32)
    assert O.lib().bho_hist_bin(np.float32(1.0), 0.0, 1.0, 256) == 255   # max -> last bin
    assert O.lib().bho_hist_bin(np.float32(-0.5), 0.0, 1.0, 256) == 0    # below -> 0
    assert O.lib().bho_hist_bin(np.float32(7.0), 0.0, 1.0, 256) == 255   # above -> last
    assert O.lib().bho_hist_bin(np.float32(0.5), 0.0, 1.0, 256) == 128
    del lim


def test_pdf_integrates_to_one():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((10000, 2)).astype(np.float32)
    h = O.histogram(a)
    width = (h.limits[:, 1] - h.limits[:, 0]) / 256
    assert np.allclose((h.pdf * width[:, None]).sum(1), 1.0, atol=1e-5)


def test_bitonic_rank_semantics():
    # distinct values: the indices by descending value (a unique answer)
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = rng.permutation(256).astype(np.float32)
        assert O.bitonic_ranks(v).astype(int).tolist() == np.argsort(-v, kind="stable").tolist()
    # ties keep the network's positional order (estimate.cl:111-113 compares
    # positions, not original indices); hand-traced through the network:
    # [1, 1, 0, 1] -> [(0,1),(1,1),(2,0),(3,1)] after length 1, the inc=2 pass
    # swaps nothing, the inc=1 pass swaps positions 2 and 3
    assert O.bitonic_ranks(np.array([1, 1, 0, 1], np.float32)).tolist() == [0, 1, 3, 2]
    for _ in range(20):
        v = rng.integers(0, 6, 256).astype(np.float32)
        r = O.bitonic_ranks(v).astype(int)
        assert sorted(r.tolist()) == list(range(256))
        assert np.all(np.diff(v[r]) <= 0)


def test_mean_variance_closed_form():
    # extensions_test.clj:18-47 style: known moments of an arithmetic sequence
    n = 1001
    a = np.stack([np.arange(n, dtype=np.float32), 2 * np.arange(n, dtype=np.float32) + 1], 1)
    mean, var = O.mean_var_f64(a)
    assert mean.tolist() == [500.0, 1001.0]
    assert var[0] == pytest.approx((n * n - 1) / 12.0, rel=1e-12)
    assert var[1] == pytest.approx(4 * (n * n - 1) / 12.0, rel=1e-12)


def test_tree_sum_shape():
    # the tree of stretch.cl:5-29 for n = 5: (((v0 + v2) + (v4 + v3)) + v1)
    v = np.array([1e8, 1.0, -1e8, 1.0, 1.0], np.float32)
    f = np.float32
    lvl1 = [f(v[0] + v[2]), f(f(v[1] + v[4]) + v[3])]  # i=5 -> 2, item 1 adds v4 first
    assert O.tree_sum_f32(v) == f(lvl1[0] + lvl1[1])
    w = np.arange(256, dtype=np.float32)
    assert O.tree_sum_f32(w) == np.float32(w.sum())
