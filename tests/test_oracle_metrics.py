"""Pins the oracle's other block metrics (itkBlockMatcher.txx:105-227) with
hand-computed answers: SSD on constant offsets and ramps, the correlation
ratio on two-bin blocks worked out with exact fractions, and each of the
correlation ratio's quirks (values <= 10 ignored, bins of <= 2 values left
out of the variance sum but kept in the totals, the 0.4 floor)."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle.oracle import pair_score

BW = 4


def _vol(block):
    """A 4x4x4 volume holding `block` (x fastest)."""
    return np.ascontiguousarray(np.asarray(block, np.float32).reshape(BW, BW, BW))


def _score(metric, base, search, bmask=None, smask=None):
    bm = None if bmask is None else np.ascontiguousarray(np.asarray(bmask, np.uint8).reshape(4, 4, 4))
    sm = None if smask is None else np.ascontiguousarray(np.asarray(smask, np.uint8).reshape(4, 4, 4))
    return pair_score(metric, _vol(base), bm, _vol(search), sm, (0, 0, 0), (0, 0, 0), BW)


def test_ssd_constant_offset_and_ramp():
    s = np.arange(64, dtype=np.float64) * 3.0
    assert _score("sd", s + 2.5, s) == 6.25
    # base = i, search = 0: mean of i^2 over 0..63 = 63*64*127/6/64
    assert _score("sd", np.arange(64), np.zeros(64)) == pytest.approx(63 * 127 / 6, rel=1e-15)


def test_ssd_masks():
    b = np.zeros(64)
    b[:10] = 4.0
    m = np.zeros(64, np.uint8)
    m[:10] = 1
    assert _score("sd", b, np.zeros(64), bmask=m) == 16.0   # only the masked 10 count
    assert _score("sd", b, np.zeros(64), smask=np.zeros(64)) == -2.0  # nothing inside


def _two_bin_block(bin0, bin1, extra_y=None):
    """Base values bin0 where search = 5 (bin 0) and bin1 where search = 15
    (bin 1); remaining voxels masked out."""
    base = np.zeros(64)
    search = np.zeros(64)
    mask = np.zeros(64, np.uint8)
    vals = list(bin0) + list(bin1)
    ys = [5.0] * len(bin0) + [15.0] * len(bin1)
    base[: len(vals)] = vals
    search[: len(vals)] = ys
    mask[: len(vals)] = 1
    return base, search, mask


def _var(xs):
    xs = [Fr(x) for x in xs]
    m = sum(xs) / len(xs)
    return sum((x - m) ** 2 for x in xs) / len(xs)


def test_cr_two_bins_exact():
    b0, b1 = [100, 102, 104, 106], [200, 204, 208, 212]
    base, search, mask = _two_bin_block(b0, b1)
    # per-bin variances 5 and 20; total variance over all eight values
    want = 1 - (_var(b0) + _var(b1)) / _var(b0 + b1)
    assert _var(b0) == 5 and _var(b1) == 20
    assert _score("cr", base, search, bmask=mask) == pytest.approx(float(want), rel=1e-13)


def test_cr_ignores_values_up_to_ten():
    b0, b1 = [100, 102, 104, 106], [200, 204, 208, 212]
    base, search, mask = _two_bin_block(b0 + [10.0, 3.0], b1 + [-50.0])
    want = 1 - (_var(b0) + _var(b1)) / _var(b0 + b1)
    assert _score("cr", base, search, bmask=mask) == pytest.approx(float(want), rel=1e-13)


def test_cr_small_bins_count_in_totals_only():
    b0, b1, b2 = [100, 102, 104, 106], [200, 204, 208, 212], [150, 160]
    base, search, mask = _two_bin_block(b0, b1)
    base[8:10] = b2
    search[8:10] = 35.0  # bin 3, two values: no variance term
    mask[8:10] = 1
    want = 1 - (_var(b0) + _var(b1)) / _var(b0 + b1 + b2)
    assert _score("cr", base, search, bmask=mask) == pytest.approx(float(want), rel=1e-13)


def test_cr_floor_and_degenerate_cases():
    # One bin: sum of variances = total variance, cr = 0 <= 0.4 -> -2.
    base, search, mask = _two_bin_block([100, 150, 120, 170, 190], [])
    assert _score("cr", base, search, bmask=mask) == -2.0
    # Identical values per bin: the variance sum is 0 -> -2.
    base, search, mask = _two_bin_block([100] * 4, [300] * 4)
    assert _score("cr", base, search, bmask=mask) == -2.0
    # Nothing inside both masks -> -2.
    assert _score("cr", base, search, bmask=np.zeros(64)) == -2.0
    # Bins follow floor(y * 0.1): y = 9.99 and y = 0 share bin 0, y = -0.01 is bin -1.
    b0, b1 = [100, 102, 104, 106], [200, 204, 208, 212]
    base, search, mask = _two_bin_block(b0, b1)
    search[:4] = [0.0, 9.99, 3.0, 9.0]
    assert _score("cr", base, search, bmask=mask) == pytest.approx(
        float(1 - (_var(b0) + _var(b1)) / _var(b0 + b1)), rel=1e-13)
    search[0] = -0.01  # moves one value of bin 0 to bin -1 (a 1-value bin)
    want = 1 - (_var(b0[1:]) + _var(b1)) / _var(b0 + b1)
    assert _score("cr", base, search, bmask=mask) == pytest.approx(float(want), rel=1e-13)


def test_cr_is_invariant_to_voxel_order():
    rng = np.random.default_rng(0)
    search = rng.uniform(0, 60, 64)
    base = 100 + 20 * search + rng.normal(0, 5, 64)
    v = _score("cr", base, search)
    assert v > 0.4
    p = rng.permutation(64)
    assert _score("cr", base[p], search[p]) == pytest.approx(v, rel=1e-12)
