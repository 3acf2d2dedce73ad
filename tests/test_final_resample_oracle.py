"""Pins the CPU restatement of the final output interpolators
(oracle/final_resample.py) against scipy.ndimage and closed forms."""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import final_resample as fr


def _rand(shape, seed=0):
    return np.random.default_rng(seed).normal(size=shape) * 100.0


@pytest.mark.parametrize("shape", [(7, 9, 11), (1, 5, 6), (2, 3, 40), (5, 24, 3)])
def test_bspline_matches_scipy_mirror(shape):
    # scipy's "mirror" mode reflects about the edge sample (d c b | a b c d),
    # the boundary ITK's B-spline interpolator uses.
    d = _rand(shape)
    rng = np.random.default_rng(1)
    nz, ny, nx = shape
    n = 500
    ci = np.stack([rng.uniform(-0.5, nx - 0.5, n), rng.uniform(-0.5, ny - 0.5, n),
                   rng.uniform(-0.5, nz - 0.5, n)], -1)
    got = fr.evaluate(d, ci, "bspline")
    ref = ndimage.map_coordinates(d, [ci[:, 2], ci[:, 1], ci[:, 0]], order=3, mode="mirror")
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-7 * np.abs(d).max())


def test_bspline_coefficients_match_scipy_filter():
    d = _rand((6, 30, 8), seed=3)  # x lines of 8 (full initial loop), y of 30 (horizon)
    ref = ndimage.spline_filter(d, order=3, mode="mirror")
    # scipy's causal initialisation is exact; ITK's stops where |z|^n < 1e-10.
    assert np.allclose(fr.bspline_coefficients(d), ref, rtol=0, atol=1e-7 * np.abs(d).max())


def test_bspline_interpolates_at_grid_points():
    d = _rand((4, 5, 6), seed=5)
    k = np.stack(np.meshgrid(np.arange(6), np.arange(5), np.arange(4), indexing="ij"), -1)
    k = k.transpose(2, 1, 0, 3).astype(float)
    assert np.allclose(fr.evaluate(d, k, "bspline"), d, atol=1e-9)


def test_nearest_and_linear_hand_values():
    d = np.arange(2 * 3 * 4, dtype=float).reshape(2, 3, 4)  # value = x + 4 y + 12 z
    ci = np.array([[0.5, 0, 0], [0.49, 1.5, 0.5], [-0.5, 0, 0], [3.49, 2, 1], [3.5, 0, 0],
                   [1.25, 0.5, 0.75]])
    nn = fr.evaluate(d, ci, "nn")
    # half up: 0.5 -> 1; (0.49, 1.5, 0.5) -> (0, 2, 1); -0.5 -> 0; 3.5 is outside.
    assert nn.tolist() == [1.0, 0 + 8 + 12, 0.0, 3 + 8 + 12, 0.0, 1 + 4 * 1 + 12 * 1]
    lin = fr.evaluate(d, ci, "linear")
    # Linear in the interior reproduces x + 4y + 12z; edges clamp.
    assert lin[0] == pytest.approx(0.5)
    assert lin[5] == pytest.approx(1.25 + 2.0 + 9.0)
    assert lin[2] == pytest.approx(0.0)
    assert lin[3] == pytest.approx(3 + 8 + 12)  # x beyond the last sample clamps


def test_sinc_delta_at_grid_points_and_half_shift_weights():
    d = _rand((12, 12, 12), seed=7)
    k = np.array([[5.0, 6.0, 4.0], [0.0, 0.0, 11.0]])
    assert np.allclose(fr.evaluate(d, k, "sinc"), [d[4, 6, 5], d[11, 0, 0]])
    # A shift of one half along x alone: weights (1 - x^2/25) sin(pi x)/(pi x)
    # at x = 0.5 - o for offsets o = -4..5, i.e. |x| = 0.5, 1.5, ..., 4.5.
    x0, y0, z0 = 5, 6, 4
    v = fr.evaluate(d, np.array([[x0 + 0.5, y0, z0]]), "sinc")[0]
    want = 0.0
    for o in range(-4, 6):
        x = 0.5 - o
        w = (1 - x * x / 25.0) * math.sin(math.pi * x) / (math.pi * x)
        want += w * d[z0, y0, min(max(x0 + o, 0), 11)]
    assert v == pytest.approx(want, rel=1e-12)
    # closed form of the nearest weight: (1 - 0.01) * 2 / pi
    assert (1 - 0.25 / 25) * math.sin(math.pi / 2) / (math.pi / 2) == pytest.approx(0.99 * 2 / math.pi)


def test_outside_is_zero_for_every_interpolator():
    d = _rand((3, 3, 3)) + 1000
    ci = np.array([[-0.51, 1, 1], [1, 2.5, 1], [1, 1, 2.5000001], [2.49, 2.49, -0.5]])
    for interp in fr.INTERP:
        v = fr.evaluate(d, ci, interp)
        assert v[:3].tolist() == [0.0, 0.0, 0.0]
        assert v[3] != 0.0


def test_index_map_identity_and_rotation():
    class G:
        def __init__(self, dims, spacing, origin, direction=tuple(np.eye(3).ravel())):
            self.dims, self.spacing, self.origin, self.direction = dims, spacing, origin, direction
    g = G((4, 3, 2), (2.0, 1.0, 0.5), (10.0, -5.0, 3.0))
    ci = fr.index_map(g, g, np.eye(3).ravel(), np.zeros(3))
    assert np.allclose(ci[1, 2, 3], (3, 2, 1))
    # A translation by one voxel along x shifts the index by one.
    ci = fr.index_map(g, g, np.eye(3).ravel(), (2.0, 0, 0))
    assert np.allclose(ci[0, 0, 0], (1, 0, 0))
