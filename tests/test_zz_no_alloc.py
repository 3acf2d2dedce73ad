"""mirorr_lts_fit_noalloc against mirorr_lts_fit (through the C ABI).

The variant without device allocation runs the same kernels on buffers
carved from its file-scope scratch array, so its fits are identical to
mirorr_lts_fit's; two calls in a row, the second smaller (the array keeps
the first call's contents), classic and symmetric, rigid and affine, and a
size above the array's limit is rejected.
"""
import numpy as np
import pytest

from mirorr_amd import phantom

pytestmark = pytest.mark.gpu


def _pairs(seed, n, affine):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-80, 80, (n, 3))
    A = phantom.euler_zxy(0.08, -0.05, 0.1)
    if affine:
        A = A @ np.diag([1.04, 0.97, 1.02])
    Y = X @ A.T + np.array([3.0, -2.0, 5.0]) + rng.normal(0, 0.5, X.shape)
    bad = rng.choice(n, int(0.3 * n), replace=False)
    Y[bad] += rng.uniform(-30, 30, (len(bad), 3))
    return X, Y


@pytest.mark.parametrize("symmetric,affine", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_noalloc_lts_equals_lts_fit(symmetric, affine):
    from mirorr_amd.registration import lts_fit
    c0 = np.array([1.0, -2.0, 0.5])
    for n in (200000, 3001):  # the second call is smaller
        X, Y = _pairs(n + symmetric + 2 * affine, n, affine)
        ref = lts_fit(bool(symmetric), bool(affine), X, Y, center=c0)
        got = lts_fit(bool(symmetric), bool(affine), X, Y, center=c0, noalloc=True)
        for a, b in zip(ref[:4], got[:4]):
            np.testing.assert_array_equal(a, b)
        assert ref[4] == got[4] and ref[4] > 0


def test_noalloc_lts_rejects_more_pairs_than_its_scratch():
    from mirorr_amd.capi import MirorrError
    from mirorr_amd.registration import lts_fit
    X = np.zeros((1048577, 3))
    with pytest.raises(MirorrError):
        lts_fit(False, False, X, X, noalloc=True)
