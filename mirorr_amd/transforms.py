"""Transform parameters in the reference's ITK conventions and error metrics.

Euler3DTransform parameters are (ax, ay, az, tx, ty, tz) with matrix
Rz Rx Ry; AffineTransform parameters are the row-major matrix then the
translation. With centre c: y = M (x - c) + c + t, i.e. offset = t + c - M c.
"""
from __future__ import annotations

import numpy as np

from .phantom import euler_zxy


def matrix_offset(params, center, affine: bool):
    params = np.asarray(params, float)
    c = np.asarray(center, float)
    if affine:
        M = params[:9].reshape(3, 3)
        t = params[9:12]
    else:
        M = euler_zxy(*params[:3])
        t = params[3:6]
    return M, t + c - M @ c


def fov_corners(vol):
    nx, ny, nz = vol.dims
    D = np.asarray(vol.direction, float).reshape(3, 3)
    s = np.asarray(vol.spacing, float)
    o = np.asarray(vol.origin, float)
    idx = np.array([[i, j, k] for i in (0, nx - 1) for j in (0, ny - 1) for k in (0, nz - 1)], float)
    return o + idx @ (D * s).T


def max_corner_distance(M1, off1, M2, off2, vol):
    """Max over the 8 field-of-view corners of |T1(p) - T2(p)| (mm)."""
    P = fov_corners(vol)
    d = (P @ np.asarray(M1).T + off1) - (P @ np.asarray(M2).T + off2)
    return float(np.linalg.norm(d, axis=1).max())


def rotation_angle_deg(M1, M2):
    """Angle (degrees) of R1 R2^T, using the polar (orthogonal) part of each."""
    def polar(m):
        u, _, vt = np.linalg.svd(np.asarray(m, float))
        return u @ vt
    R = polar(M1) @ polar(M2).T
    c = np.clip((np.trace(R) - 1.0) / 2.0, -1.0, 1.0)
    return float(np.degrees(np.arccos(c)))
