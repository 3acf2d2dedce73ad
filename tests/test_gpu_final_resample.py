"""Parity of the device final-output resampler (mirorr_resample_final) with
the CPU restatement in oracle/final_resample.py, every interpolator."""
import numpy as np
import pytest

from mirorr_amd import phantom
from mirorr_amd.capi import Geom as CGeom
from oracle import final_resample as fr

pytestmark = pytest.mark.gpu


class _G:
    def __init__(self, dims, spacing, origin, direction=tuple(np.eye(3).ravel())):
        self.dims, self.spacing, self.origin, self.direction = dims, spacing, origin, direction


CASES = [
    # (input dims, output dims, direction of the input grid)
    ((28, 24, 20), (30, 26, 18), tuple(np.eye(3).ravel())),
    ((17, 2, 9), (15, 4, 11), tuple(np.eye(3).ravel())),           # a two-sample axis
    ((12, 16, 1), (14, 14, 1), tuple(np.eye(3).ravel())),          # a single slice
    ((20, 18, 16), (20, 18, 16), (0, 1, 0, 1, 0, 0, 0, 0, 1.0)),    # ARI grid
]


@pytest.mark.parametrize("interp", ["nn", "linear", "bspline", "sinc"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_final_resample_matches_oracle(interp, case):
    from mirorr_amd.registration import resample_final
    din, dout, dirn = CASES[case]
    rng = np.random.default_rng(case)
    a = rng.uniform(0, 1000, din[::-1]).astype(np.float32)
    spi, org = (1.5, 1.2, 2.0), (-20.0, -14.0, -19.0)
    gi = _G(din, spi, org, dirn)
    go = _G(dout, (1.4, 1.3, 2.1), (-21.0, -16.0, -19.0 if din[2] == 1 else -18.0))
    M = phantom.euler_zxy(0.1, -0.05, 0.15)
    if din[2] == 1:
        M = phantom.euler_zxy(0.0, 0.0, 0.15)
    off = np.array([1.3, -2.2, 0.0 if din[2] == 1 else 0.7])
    print("final_resample", interp, din, dout, flush=True)
    got = resample_final(a, CGeom.make(din, spi, org, dirn), M, off,
                         CGeom.make(dout, go.spacing, go.origin), interp)
    ref = fr.resample_final(a, gi, M.ravel(), off, go, interp)
    # Points within 1e-9 of a cell edge or of the buffer bounds may be
    # classified differently (device fma vs numpy rounding); allow a few.
    bad = ~np.isclose(got, ref, rtol=1e-5, atol=1e-2)
    assert bad.mean() < 2e-3, (bad.sum(), np.abs(got - ref).max())
    assert np.count_nonzero(got) > 0.05 * got.size


def test_final_resample_identity_grid_points():
    """On the input grid itself every interpolator returns the data (the
    B-spline interpolates, the sinc weights are a delta)."""
    from mirorr_amd.registration import resample_final
    rng = np.random.default_rng(9)
    a = rng.uniform(-500, 500, (9, 10, 11)).astype(np.float32)
    g = CGeom.make((11, 10, 9), (1.0, 2.0, 0.5), (3.0, -4.0, 5.0))
    for interp in ("nn", "linear", "bspline", "sinc"):
        got = resample_final(a, g, np.eye(3), np.zeros(3), g, interp)
        np.testing.assert_allclose(got, a, rtol=1e-5, atol=1e-3, err_msg=interp)
