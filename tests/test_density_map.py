"""The 64 x 64 -> 96 x 96 map of a density perturbation (trig_interp_matrix),
pinned by what trigonometric interpolation must satisfy: band-limited modes
|k| < n/2 are reproduced exactly at the new points, the Nyquist mode becomes
its cosine (its sine part vanishes on the coarse samples), rows at points the
two grids share are unit vectors, n -> n is the identity, and interpolating
up then sampling back at the coarse points returns the input."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("n,m", [(64, 96), (16, 24), (10, 25)])
def test_band_limited_modes_are_exact(vlp, n, m):
    P = vlp.trig_interp_matrix(n, m)
    xc, xf = np.arange(n) / n, np.arange(m) / m
    for k in range(n // 2):
        for fn in (np.cos, np.sin):
            np.testing.assert_allclose(P @ fn(2 * math.pi * k * xc), fn(2 * math.pi * k * xf), atol=1e-12)
    np.testing.assert_allclose(P @ np.cos(math.pi * n * xc), np.cos(math.pi * n * xf), atol=1e-12)


def test_shared_points_and_identity(vlp):
    P = vlp.trig_interp_matrix(64, 96)
    for i in range(0, 64, 2):          # x = i/64 = (3i/2)/96
        row = P[3 * i // 2]
        assert abs(row[i] - 1) < 1e-13 and np.max(np.abs(np.delete(row, i))) < 1e-13
    np.testing.assert_allclose(vlp.trig_interp_matrix(12, 12), np.eye(12), atol=1e-13)


def test_up_then_back_is_identity(vlp):
    """Sampling the fine interpolant at the coarse points (the fine -> coarse
    map of a band-limited field) returns the coarse data."""
    rng = np.random.default_rng(5)
    d = rng.standard_normal((64, 64))
    P = vlp.trig_interp_matrix(64, 96)
    fine = P @ d @ P.T
    back = vlp.trig_interp_matrix(96, 64) @ fine @ vlp.trig_interp_matrix(96, 64).T
    np.testing.assert_allclose(back, d, atol=1e-11)
