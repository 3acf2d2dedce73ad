"""CPU restatement of the final output resampling (test infrastructure only).

Only tests/ may import this module. It restates, in numpy and double
precision, what the reference's GetResampledImage does with the interpolator
GetInterpolatorFromString picks (itkMirorrPyramidImplement.txx:364-428,
436-470): ITK ResampleImageFilter (default value 0; a point is evaluated only
when its continuous index lies in [-0.5, n-0.5) on every axis) with
  "nn"       NearestNeighborInterpolateImageFunction (round half up),
  "linear"   LinearInterpolateImageFunction (neighbours clamped),
  "bspline"  BSplineInterpolateImageFunction of order 3: coefficients by
             Unser's recursive prefilter (IEEE Signal Processing Magazine
             16(6), 1999, Box 2; pole sqrt(3)-2, mirror boundaries, causal
             initialisation truncated at |z|^n < 1e-10), evaluation over the
             4 nearest coefficients per axis with mirrored indices,
  "sinc"     WindowedSincInterpolateImageFunction<radius 5, Welch window>
             with zero-flux Neumann edges (offsets -4..5 about floor).
ITK is a dependency the reference does not vendor; the version is not
pinned by its build files.

Pinning (tests/test_final_resample_oracle.py): the B-spline path against
scipy.ndimage (spline_filter + map_coordinates, order 3, mode "mirror"), and
the interpolation property at grid points; nearest neighbour and linear
against hand-computed values; the sinc path against the closed-form
Welch-sinc weights at half-integer shifts and the delta at grid points.
"""
from __future__ import annotations

import math

import numpy as np

INTERP = {"nn": 0, "linear": 1, "bspline": 2, "sinc": 3}


def index_map(gin, gout, matrix, offset):
    """Continuous input index of every output voxel, shape (nz, ny, nx, 3)
    in (x, y, z) order: ci = S_in^-1 D_in^-1 (M p + off - o_in), p the
    physical point of the output voxel."""
    nx, ny, nz = gout.dims
    k = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    k = k.transpose(2, 1, 0, 3).astype(np.float64)
    Dout = np.asarray(gout.direction, float).reshape(3, 3)
    Din = np.asarray(gin.direction, float).reshape(3, 3)
    p = np.asarray(gout.origin, float) + (k * np.asarray(gout.spacing, float)) @ Dout.T
    q = p @ np.asarray(matrix, float).reshape(3, 3).T + np.asarray(offset, float)
    u = (q - np.asarray(gin.origin, float)) @ np.linalg.inv(Din).T
    return u / np.asarray(gin.spacing, float)


def bspline_coefficients(data):
    """Cubic B-spline coefficients (x, y, z passes, double)."""
    c = np.asarray(data, np.float64).copy()
    z = math.sqrt(3.0) - 2.0
    gain = (1.0 - z) * (1.0 - 1.0 / z)
    horizon = math.ceil(math.log(1e-10) / math.log(abs(z)))
    for axis in (2, 1, 0):  # numpy axes of x, y, z
        N = c.shape[axis]
        if N == 1:
            continue
        a = np.moveaxis(c, axis, 0)  # view: a[n] is element n of every line
        a *= gain
        if horizon < N:
            s = a[0].copy()
            zn = z
            for n in range(1, horizon):
                s += zn * a[n]
                zn *= z
            a[0] = s
        else:
            iz = 1.0 / z
            zn = z
            z2n = z ** (N - 1)
            s = a[0] + z2n * a[N - 1]
            z2n *= z2n * iz
            for n in range(1, N - 1):
                s += (zn + z2n) * a[n]
                zn *= z
                z2n *= iz
            a[0] = s / (1.0 - zn * zn)
        for n in range(1, N):
            a[n] += z * a[n - 1]
        a[N - 1] = (z / (z * z - 1.0)) * (z * a[N - 2] + a[N - 1])
        for n in range(N - 2, -1, -1):
            a[n] = z * (a[n + 1] - a[n])
    return c


def _mirror(i, N):
    """Whole-sample symmetric index for i in -2..N+1 (two reflections cover
    N = 2, where ITK's single reflection would leave its buffer)."""
    if N == 1:
        return np.zeros_like(i)
    for _ in range(2):
        i = np.where(i < 0, -i, i)
        i = np.where(i > N - 1, 2 * (N - 1) - i, i)
    return i


def _cubic(w):
    return [(1 - w) ** 3 / 6, (4 - 6 * w ** 2 + 3 * w ** 3) / 6,
            (1 + 3 * w + 3 * w ** 2 - 3 * w ** 3) / 6, w ** 3 / 6]


def _sinc_w(x):
    px = np.pi * x
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(x == 0.0, 1.0, np.sin(px) / px)
    return (1.0 - x * x / 25.0) * s


def evaluate(data, ci, interp, coef=None):
    """Interpolates data (nz, ny, nx) at continuous indices ci (..., 3 in
    x, y, z order); 0 outside [-0.5, n-0.5)."""
    data = np.asarray(data, np.float64)
    nz, ny, nx = data.shape
    dims = (nx, ny, nz)
    cx, cy, cz = ci[..., 0], ci[..., 1], ci[..., 2]
    inside = np.ones(cx.shape, bool)
    for c, n in zip((cx, cy, cz), dims):
        inside &= (c >= -0.5) & (c < n - 0.5)
    cx, cy, cz = (np.where(inside, c, 0.0) for c in (cx, cy, cz))
    if interp == "nn":
        ix, iy, iz = (np.minimum(np.floor(c + 0.5).astype(int), n - 1)
                      for c, n in zip((cx, cy, cz), dims))
        v = data[iz, iy, ix]
    elif interp == "linear":
        b = [np.floor(c).astype(int) for c in (cx, cy, cz)]
        f = [c - bb for c, bb in zip((cx, cy, cz), b)]
        for k in range(3):
            neg = b[k] < 0
            b[k] = np.where(neg, 0, b[k])
            f[k] = np.where(neg, 0.0, f[k])
        hi = [np.minimum(b[k] + 1, dims[k] - 1) for k in range(3)]
        v = np.zeros(cx.shape)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = ((f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) *
                         (f[2] if dz else 1 - f[2]))
                    v += w * data[hi[2] if dz else b[2], hi[1] if dy else b[1],
                                  hi[0] if dx else b[0]]
    elif interp == "bspline":
        c = bspline_coefficients(data) if coef is None else coef
        b = [np.floor(t).astype(int) for t in (cx, cy, cz)]
        w = [_cubic(t - bb) for t, bb in zip((cx, cy, cz), b)]
        v = np.zeros(cx.shape)
        for kz in range(4):
            iz = _mirror(b[2] - 1 + kz, nz)
            for ky in range(4):
                iy = _mirror(b[1] - 1 + ky, ny)
                for kx in range(4):
                    ix = _mirror(b[0] - 1 + kx, nx)
                    v += w[0][kx] * w[1][ky] * w[2][kz] * c[iz, iy, ix]
    elif interp == "sinc":
        b = [np.floor(t).astype(int) for t in (cx, cy, cz)]
        d = [t - bb for t, bb in zip((cx, cy, cz), b)]
        offs = range(-4, 6)
        W = [[np.where(dd == 0.0, 1.0 if o == 0 else 0.0, _sinc_w(dd - o)) for o in offs]
             for dd in d]
        v = np.zeros(cx.shape)
        for a, oz in enumerate(offs):
            iz = np.clip(b[2] + oz, 0, nz - 1)
            for bq, oy in enumerate(offs):
                iy = np.clip(b[1] + oy, 0, ny - 1)
                for cq, ox in enumerate(offs):
                    ix = np.clip(b[0] + ox, 0, nx - 1)
                    v += W[0][cq] * W[1][bq] * W[2][a] * data[iz, iy, ix]
    else:
        raise ValueError(interp)
    return np.where(inside, v, 0.0)


def resample_final(data, gin, matrix, offset, gout, interp):
    """out(p) = data(matrix p + offset) on the output grid, float32."""
    ci = index_map(gin, gout, matrix, offset)
    return evaluate(data, ci, interp).astype(np.float32)
