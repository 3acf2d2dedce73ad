"""Binary16 volume storage: the host encoding (mirorr_half_encode), CPU only.

The encoding is (v - offset) * scale with offset the nearest integer to the
middle of the finite range and scale a power of two bringing the larger of
the half-range and |offset| into (1024, 2048]: integer intensities of up to
12 bits, and the encoding of 0, are exact. Checked against numpy's own
float32 -> float16 rounding (round to nearest even)."""
import ctypes as C

import numpy as np
import pytest

from mirorr_amd import capi


def encode(data):
    data = np.ascontiguousarray(data, np.float32)
    out = np.zeros(data.size, np.uint16)
    off, sc = C.c_double(), C.c_double()
    st = capi.lib().mirorr_half_encode(capi.ptr(data), data.size, capi.ptr(out), C.byref(off),
                                       C.byref(sc))
    assert st == capi.OK
    return out.view(np.float16).reshape(data.shape), off.value, sc.value


@pytest.mark.parametrize("lo,hi", [(0, 4095), (-1024, 3071), (-2048, 2047), (0, 255), (0, 1)])
def test_integer_intensities_are_exact(lo, hi):
    rng = np.random.default_rng(abs(lo) + 7)
    d = rng.integers(lo, hi + 1, size=20000).astype(np.float32)
    d[:2] = [lo, hi]
    h, off, sc = encode(d)
    assert off == np.rint(0.5 * (lo + hi)) and sc == 2.0 ** np.floor(np.log2(sc))
    np.testing.assert_array_equal(h.astype(np.float64) / sc + off, d)
    # the encoding of intensity 0 (written outside the field of view) is exact
    z = np.float16((0.0 - off) * sc)
    assert float(z) / sc + off == 0.0


def test_float_intensities_round_to_nearest():
    rng = np.random.default_rng(3)
    d = (rng.normal(300.0, 80.0, 50000)).astype(np.float32)
    d[0] = np.float32(np.nan)
    h, off, sc = encode(d)
    fin = np.isfinite(d)
    want = ((d[fin] - np.float32(off)) * np.float32(sc)).astype(np.float16)
    np.testing.assert_array_equal(h[fin].view(np.uint16), want.view(np.uint16))
    assert np.isnan(h[0])
    # largest encoded magnitude within (1024, 2048] (the half-range or |offset|)
    m = max(np.abs(h[fin].astype(np.float64)).max(), abs(off * sc))
    assert 1024.0 < m <= 2048.0
    # relative rounding of binary16: half an ulp of the magnitude
    err = np.abs(h[fin].astype(np.float64) / sc + off - d[fin])
    assert err.max() <= 2.0 ** -11 * m / sc


def test_constant_and_empty():
    h, off, sc = encode(np.full(10, 5.0, np.float32))
    assert off == 5.0 and np.all(h == 0)
    h, off, sc = encode(np.zeros(0, np.float32))
    assert off == 0.0 and sc == 1.0
