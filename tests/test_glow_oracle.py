"""Pins of the glow-curve oracle (oracle/glow.py, plots.py:20-66) and the host
side of the time-axis side-car.

The oracle calls np.histogram / np.convolve as the reference does; these tests
fix what those calls compute with independent loops and closed forms, so that a
changed edge rule (half-open bins, closed last bin, np.arange edge count), a
dropped division or a wrong window fails here.
"""
import math

import numpy as np
import pytest

from oracle import glow


def brute_hist(T, w, bw):
    """Loop-by-loop weighted histogram over edges j * bw, j < ceil((max + bw) / bw)."""
    n_edges = int(math.ceil((max(T) + bw) / bw))
    edges = [j * bw for j in range(n_edges)]
    h = [0.0] * (n_edges - 1)
    for x, wi in zip(T, w):
        for j in range(n_edges - 1):
            last = j == n_edges - 2
            if edges[j] <= x and (x < edges[j + 1] or (last and x <= edges[j + 1])):
                h[j] += wi
                break
    return np.array(h) / bw


def brute_smooth(h, k):
    if len(h) >= k:
        return np.array([sum(h[m:m + k]) / k for m in range(len(h) - k + 1)])
    return np.array([sum(h) / k] * (k - len(h) + 1))


@pytest.mark.parametrize("bw,win", [(1.0, 50.0), (0.5, 3.0), (2.5, 10.0), (1.0, 0.3)])
def test_hist_and_smooth_matches_brute_force(bw, win):
    rng = np.random.default_rng(11)
    for _ in range(5):
        T = rng.uniform(0, 120, size=300)
        T[:20] = np.round(T[:20])               # values on edges
        T[20] = np.floor(T.max() / bw) * bw + bw  # a value that becomes the max
        w = rng.integers(0, 5, size=T.size) / 3.0
        got = glow.hist_and_smooth(T, w, bw, win)
        k = max(1, int(win / bw))
        want = brute_smooth(brute_hist(T, w, bw), k)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_last_edge_is_closed_and_negative_values_drop():
    # max T = 3.0 exactly: edges 0,1,2,3 (ceil(4.0) = 4 edges); 3.0 lands in the last bin
    T = np.array([-0.5, 0.0, 1.0, 2.5, 3.0])
    w = np.array([7.0, 1.0, 2.0, 3.0, 4.0])
    h = glow.hist_and_smooth(T, w, 1.0, 0.0)        # k = 1: the histogram itself
    np.testing.assert_array_equal(h, [1.0, 2.0, 7.0])


def test_constant_rate_closed_form():
    # one unit of weight at the centre of every bin: histogram 1 everywhere, smooth 1
    T = np.arange(0.5, 400.5, 1.0)
    s = glow.hist_and_smooth(T, np.ones_like(T), 1.0, 50.0)
    assert s.shape == (400 - 50 + 1,)
    np.testing.assert_allclose(s, 1.0, rtol=1e-13)


def test_single_event_window():
    # weight 5 at T = 70.2 -> bin 70 of 71 bins (edges 0..71); k = 10 windows m = 61..70 see 5/10
    s = glow.hist_and_smooth(np.array([70.2]), np.array([5.0]), 1.0, 10.0)
    assert s.shape == (71 - 10 + 1,)
    want = np.zeros(62)
    want[61:71] = 0.5
    np.testing.assert_allclose(s, want, atol=1e-15)


def test_window_wider_than_histogram():
    # 3 bins, k = 10: np.convolve swaps its operands; every point is total / k
    s = glow.hist_and_smooth(np.array([0.5, 1.5, 2.5]), np.array([1.0, 2.0, 3.0]), 1.0, 10.0)
    np.testing.assert_allclose(s, np.full(8, 0.6), rtol=1e-14)


def test_weight_is_conserved_by_the_histogram():
    rng = np.random.default_rng(3)
    T = rng.uniform(0, 500, 2000)
    w = rng.uniform(0, 2, 2000)
    h = glow.hist_and_smooth(T, w, 2.0, 0.0)
    assert h.sum() * 2.0 == pytest.approx(w.sum(), rel=1e-12)


def test_running_mean_of_a_line():
    a = np.arange(20, dtype=np.float64)
    np.testing.assert_allclose(glow.running_mean(a, 5), np.arange(2, 18), rtol=1e-14)


def test_tl_curves_uses_first_sim_axis_and_mask():
    # steps 4, sims 2, runs 1; the second sim's axis differs and is ignored; step 3 is padding
    x_ax = np.zeros((4, 2, 1))
    x_ax[:3, 0, 0] = [1.0, 2.0, 3.0]
    x_ax[:3, 1, 0] = [9.0, 9.0, 9.0]
    lum = np.zeros((4, 2, 1))
    lum[:, 0, 0] = [1, 0, 2, 5]
    lum[:, 1, 0] = [1, 2, 0, 5]
    (c,) = glow.tl_curves(x_ax, lum, [(10.0, 2.0)], 1.0, 0.0)
    # T = 12, 14, 16 -> edges 0..16 (16 bins, the last one [15, 16] closed): bins 12, 14
    # and 15 carry the sim means 1, 1, 1
    want = np.zeros(16)
    want[[12, 14, 15]] = 1.0
    np.testing.assert_allclose(c, want)


def test_x_ax_side_car_round_trip(tmp_path):
    from mcluminescence_amd.simulate import (read_results, write_results_csv, write_x_ax_csv,
                                             x_ax_path)
    rng = np.random.default_rng(5)
    lum = rng.integers(0, 3, size=(6, 3, 2)).astype(np.float64)
    er = rng.uniform(0, 1, size=lum.shape)
    x = np.cumsum(rng.uniform(0.1, 1.0, size=lum.shape), axis=0)
    main = write_results_csv(tmp_path / "exp_rt.csv", lum, er)
    side = write_x_ax_csv(x_ax_path(main), x)
    assert side.name == "exp_rt_x_ax.csv"
    assert side.read_text().splitlines()[0] == "run,sim,step,t"
    x2, lum2, er2 = read_results("rt", results_dir=tmp_path)
    np.testing.assert_array_equal(lum2, lum)
    np.testing.assert_array_equal(er2, er)
    np.testing.assert_array_equal(x2, x)
