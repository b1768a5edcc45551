"""Bad-projection detection (find_bad_ele.py steps) on the CPU."""
import math

import numpy as np
import pytest

from graptor_amd import badprojections as bp


def test_reduce_noise_against_loops():
    rng = np.random.default_rng(0)
    U = rng.standard_normal((9, 11))
    s = 2
    got = bp.reduce_noise(U, s)
    ref = np.zeros((9 - 2 * s, 11 - 2 * s))
    for i in range(ref.shape[0]):
        for j in range(ref.shape[1]):
            ref[i, j] = U[i:i + 2 * s + 1, j:j + 2 * s + 1].sum() / (4 * s * s)
    assert np.allclose(got, ref)


def test_compare_shift_finds_shifted_copy():
    rng = np.random.default_rng(1)
    U = rng.standard_normal((20, 20))
    V = np.roll(U, (1, -2), axis=(0, 1))
    n, E = bp.compare_shift(U, V, 3)
    assert n == 0.0 and E.shape == (14, 14)
    n, _ = bp.compare_shift(U, V, 1)  # shift (1, -2) outside radius 1
    assert n > 0
    n2, E2 = bp.compare_shift_twosided(U, U, 2)
    assert n2 == 0.0


def test_weighted_gradient_against_loops():
    rng = np.random.default_rng(2)
    u = rng.uniform(0.5, 2.0, (6, 7))
    _, got = bp.weighted_gradient(u)
    ref = np.zeros((4, 5))
    for x in range(1, 5):
        for y in range(1, 6):
            slopes, nbs = [], []
            for i in (-1, 0, 1):
                for j in (-1, 0, 1):
                    if i == 0 and j == 0:
                        continue
                    slopes.append(abs(u[x, y] - u[x + i, y + j]) / math.hypot(i, j))
                    nbs.append(abs(u[x + i, y + j]))
            ref[x - 1, y - 1] = max(slopes) * min(nbs)
    assert np.allclose(got, ref)
    _, flat = bp.weighted_gradient(np.ones((5, 5)))
    assert (flat == 0).all()


def test_combined_errors_and_selection():
    e1 = [1.0, 5.0, 1.0, 1.0]       # pairs (0,1) (1,2) (2,3) (3,4)
    e2 = [0.0, 0.0, 0.0]            # pairs (0,2) (1,3) (2,4)
    comb = bp.combined_errors(e1, e2)
    # projection i: min(e1[i] + e2[i]/2, e1[i-1] + e2[i-2]/2), clipped indices
    assert comb == [1.0, 1.0, 1.0, 1.0, 1.0]
    e1 = [1.0, 5.0, 5.0, 1.0]
    comb = bp.combined_errors(e1, e2)
    assert comb[2] == 5.0 and max(comb) == 5.0
    assert bp.select_worst([1.0, 3.0, 3.0, 2.0], 0.5) == [2, 3]
    assert bp.select_worst([1.0, 3.0, 3.0, 2.0], 3) == [2, 3, 4]
    assert bp.select_worst([1.0, 3.0, 2.0, 0.5], 0.125) == [2]   # round(0.5) -> 1


def test_detects_a_corrupted_projection():
    rng = np.random.default_rng(3)
    NA, Nz, Nx = 12, 40, 48
    yy, xx = np.mgrid[0:Nz, 0:Nx]
    series = []
    for k in range(NA):
        img = np.zeros((Nz, Nx)) + 0.2
        for cy, cx, r in ((15, 14 + k, 6), (28, 30 - k // 2, 5)):
            img[(yy - cy) ** 2 + (xx - cx) ** 2 < r * r] += 1.0
        series.append(img)
    series = np.stack(series)
    series[7] = np.roll(series[7], 9, axis=0) + 0.5 * rng.uniform(size=(Nz, Nx))
    bad = bp.find_bad_projections_array(bp.brightness_rescale(series), searchradius=3,
                                        fraction=1)
    assert bad == [8]  # 1-based
