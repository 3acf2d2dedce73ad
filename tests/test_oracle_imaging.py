"""Pins the oracle's resampler, Gaussian pyramid and schedule tuner
(ITK filters restated; itkPyramidScheduleTuner.txx:299-414) with invariants,
closed forms and known answers."""
import numpy as np
import pytest
from scipy.special import ive

from oracle.oracle import Geom

I3 = np.eye(3).ravel()


def _rand(shape, seed):
    return np.random.default_rng(seed).uniform(0, 100, shape).astype(np.float32)


def test_identity_resample_is_a_copy(oracle):
    a = _rand((7, 9, 11), 0)
    g = Geom(dims=a.shape[::-1], spacing=(1.5, 2.0, 0.7), origin=(-3.0, 1.0, 5.0))
    out, m = oracle.resample(a, None, g, I3, [0, 0, 0], g, want_mask=True)
    np.testing.assert_allclose(out, a, rtol=1e-6)
    assert np.all(m == 255)


def test_integer_translation_is_exact_shift_and_outside_is_zero(oracle):
    a = _rand((8, 10, 12), 1)
    g = Geom(dims=a.shape[::-1], spacing=(2.0, 1.0, 3.0))
    # y = x + (4, -2, 3) mm = (+2, -2, +1) voxels
    out, m = oracle.resample(a, None, g, I3, [4.0, -2.0, 3.0], g, want_mask=True)
    np.testing.assert_allclose(out[:-1, 2:, :-2], a[1:, :-2, 2:], rtol=1e-6)
    # beyond half a voxel outside the input: default value 0, mask 0
    assert np.all(out[:, :, -2:] == 0) and np.all(m[:, :, -2:] == 0)
    assert np.all(m[:-1, 2:, :-2] == 255)


def test_trilinear_reproduces_linear_fields(oracle):
    nz, ny, nx = 10, 11, 12
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = (3.0 * x - 2.0 * y + 0.5 * z + 7.0).astype(np.float32)
    g = Geom(dims=(nx, ny, nz))
    ang = 0.1
    M = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    off = np.array([0.3, 0.2, -0.1])
    out = oracle.resample(f, None, g, M.ravel(), off, g)
    q = np.stack([x, y, z], -1) @ M.T + off
    inside = np.all((q >= 0) & (q <= np.array([nx - 1, ny - 1, nz - 1])), -1)
    expect = 3.0 * q[..., 0] - 2.0 * q[..., 1] + 0.5 * q[..., 2] + 7.0
    np.testing.assert_allclose(out[inside], expect[inside], atol=1e-4)


def test_half_voxel_border_clamps_to_edge(oracle):
    a = _rand((4, 4, 4), 2)
    g = Geom(dims=(4, 4, 4))
    out = oracle.resample(a, None, g, I3, [-0.4, 0, 0], g)  # x' = x - 0.4
    np.testing.assert_allclose(out[:, :, 0], a[:, :, 0], rtol=1e-6)  # index -0.4 -> edge
    out = oracle.resample(a, None, g, I3, [-0.6, 0, 0], g)
    assert np.all(out[:, :, 0] == 0)  # index -0.6 is outside


def test_mask_truncation(oracle):
    m = np.zeros((2, 2, 4), np.uint8)
    m[:, :, 2:] = 255
    g = Geom(dims=(4, 2, 2))
    # sample halfway between voxel 1 (0) and 2 (255): 127.5 -> 127
    _, om = oracle.resample(np.zeros_like(m, dtype=np.float32), m, g, I3, [0.5, 0, 0], g,
                            want_mask=True)
    assert om[0, 0, 1] == 127


@pytest.mark.parametrize("var", [0.0, 1.0, 4.0, 16.0])
def test_gaussian_kernel_matches_bessel(oracle, var):
    k = oracle.gaussian_kernel(var)
    full = np.concatenate([k[:0:-1], k])
    assert abs(full.sum() - 1.0) < 1e-12
    if var == 0.0:
        np.testing.assert_allclose(full, [0, 1, 0], atol=0)
        return
    # e^-t I_n(t) normalised over the truncated support (A&S approximations
    # agree with the exact Bessel values to ~1e-7)
    n = np.arange(len(k))
    exact = ive(n, var)
    exact = exact / (exact[0] + 2 * exact[1:].sum())
    np.testing.assert_allclose(k, exact, atol=2e-7)
    # the truncation keeps at least 1 - max_error of the mass
    assert ive(0, var) + 2 * ive(n[1:], var).sum() >= 0.9


def test_gaussian_smoothing_constant_and_separable(oracle):
    c = np.full((9, 10, 11), 42.0, np.float32)
    np.testing.assert_allclose(oracle.gaussian_smooth(c, [1.0, 4.0, 1.0]), c, rtol=1e-6)
    imp = np.zeros((15, 15, 15), np.float32)
    imp[7, 7, 7] = 1.0
    out = oracle.gaussian_smooth(imp, [1.0, 4.0, 0.0])
    kx = oracle.gaussian_kernel(1.0)
    ky = oracle.gaussian_kernel(4.0)
    fx = np.zeros(15)
    fy = np.zeros(15)
    for i, v in enumerate(kx):
        fx[7 + i] = fx[7 - i] = v
    for i, v in enumerate(ky):
        fy[7 + i] = fy[7 - i] = v
    expect = np.zeros((15, 15, 15))
    expect[7] = np.outer(fy, fx)
    np.testing.assert_allclose(out, expect, atol=1e-7)


def test_pyramid_level_geometry(oracle):
    g = Geom(dims=(256, 255, 128), spacing=(1.0, 1.2, 2.0), origin=(-10.0, 0.0, 5.0))
    lg = oracle.pyramid_level_geom(g, (2, 2, 4))
    assert tuple(lg.dims) == (128, 127, 32)
    np.testing.assert_allclose(lg.spacing, (2.0, 2.4, 8.0))
    np.testing.assert_allclose(lg.origin, (-10.0 + 0.5, 0.0 + 0.6, 5.0 + 3.0))


def test_pyramid_step_of_constant(oracle):
    c = np.full((16, 20, 24), 5.0, np.float32)
    g = Geom(dims=(24, 20, 16))
    lg = oracle.pyramid_level_geom(g, (2, 2, 2))
    out = oracle.pyramid_step(c, g, (2, 2, 2), lg)
    assert out.shape == (8, 10, 12)
    np.testing.assert_allclose(out, 5.0, rtol=1e-6)


def test_schedule_known_answers(oracle):
    iso = lambda s: (s, s, s)
    m, f = oracle.schedule_symmetric((128, 128, 64), iso(1.5), (128, 128, 64), iso(1.5))
    assert m == f == [[4, 4, 2], [2, 2, 2], [1, 1, 1]]
    m, _ = oracle.schedule_symmetric((128, 128, 64), iso(1.5), (128, 128, 64), iso(1.5),
                                     max_levels=2)
    assert m == [[2, 2, 2], [1, 1, 1]]
    m, _ = oracle.schedule_symmetric((256, 256, 128), iso(1.0), (256, 256, 128), iso(1.0))
    assert m == [[8, 8, 4], [4, 4, 4], [2, 2, 2], [1, 1, 1]]
    m, _ = oracle.schedule_symmetric((512, 512, 256), iso(1.0), (512, 512, 256), iso(1.0))
    assert m == [[16, 16, 8], [8, 8, 8], [4, 4, 4], [2, 2, 2], [1, 1, 1]]


def test_schedule_aligns_anisotropic_resolutions(oracle):
    # moving 1 mm, fixed 3 mm in z: the moving image is first shrunk alone in z
    m, f = oracle.schedule_symmetric((256, 256, 256), (1.0, 1.0, 1.0), (256, 256, 86),
                                     (1.0, 1.0, 3.0))
    assert m[-1] == [1, 1, 1] and f[-1] == [1, 1, 1]
    assert m[-2][2] == 2 and f[-2][2] == 1  # crude resolution alignment in z first
