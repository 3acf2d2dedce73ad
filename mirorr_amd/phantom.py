"""Seeded synthetic volume pairs for the registration workloads.

A phantom is a sum of randomly oriented solid ellipsoids (semi-axes and
intensities drawn uniformly) plus Gaussian noise. The moving volume is the
same phantom under a contrast remap I' = 1000 - 0.8 I + 0.0005 I^2, evaluated
analytically at the transformed coordinates (no interpolation blur), with
fresh noise. Volumes are float32 numpy arrays of shape (nz, ny, nx) whose
grid is centred on the physical origin with identity direction.

The ground-truth transform maps moving physical points to fixed physical
points (the ITK convention of the reference: Fixed(T(x)) ~ Moving(x)).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Volume:
    data: np.ndarray          # float32, (nz, ny, nx)
    spacing: tuple            # (sx, sy, sz) mm
    origin: tuple             # (ox, oy, oz) mm
    direction: tuple = tuple(np.eye(3).ravel())
    mask: np.ndarray | None = None  # uint8 (nz, ny, nx) or None (= all 255)

    @property
    def dims(self):
        return tuple(int(v) for v in self.data.shape[::-1])


def centred_origin(dims, spacing):
    return tuple(-0.5 * (d - 1) * s for d, s in zip(dims, spacing))


def euler_zxy(ax, ay, az):
    """Rotation matrix of ITK's Euler3DTransform default (Z*X*Y) order."""
    cx, sx = math.cos(ax), math.sin(ax)
    cy, sy = math.cos(ay), math.sin(ay)
    cz, sz = math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ rx @ ry


def random_rigid(rng, max_deg=10.0, max_mm=10.0):
    ang = np.deg2rad(rng.uniform(-max_deg, max_deg, 3))
    t = rng.uniform(-max_mm, max_mm, 3)
    return euler_zxy(*ang), t, ang


def random_affine(rng, max_deg=10.0, scale=(0.9, 1.1), shear=0.05, max_mm=15.0):
    r, _, ang = random_rigid(rng, max_deg, 0.0)
    s = np.diag(rng.uniform(scale[0], scale[1], 3))
    sh = np.eye(3)
    sh[0, 1], sh[0, 2], sh[1, 2] = rng.uniform(-shear, shear, 3)
    t = rng.uniform(-max_mm, max_mm, 3)
    return r @ sh @ s, t


def _ellipsoids(rng, n, axes_mm=(5.0, 40.0), intens=(0.0, 1000.0), extent_mm=None):
    ells = []
    for _ in range(n):
        axes = rng.uniform(axes_mm[0], axes_mm[1], 3)
        centre = np.array([rng.uniform(-0.4 * e, 0.4 * e) for e in extent_mm])
        rot = euler_zxy(*rng.uniform(-math.pi, math.pi, 3))
        val = rng.uniform(intens[0], intens[1])
        ells.append((centre, axes, rot, val))
    return ells


def _render(ells, dims, spacing, origin, A=None, t=None):
    """Sum of ellipsoid indicator functions evaluated at y = A x + t for every
    voxel centre x of the grid (A = identity when None)."""
    nx, ny, nz = dims
    out = np.zeros((nz, ny, nx), np.float32)
    sp = np.asarray(spacing, float)
    org = np.asarray(origin, float)
    if A is None:
        A = np.eye(3)
        t = np.zeros(3)
    Ainv = np.linalg.inv(A)
    for centre, axes, rot, val in ells:
        # bounding box of the ellipsoid in y space -> corners mapped to x space
        r = np.abs(rot) @ axes
        corners = np.array([[centre[0] + sx * r[0], centre[1] + sy * r[1], centre[2] + sz * r[2]]
                            for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        xc = (corners - t) @ Ainv.T
        lo = np.floor((xc.min(0) - org) / sp).astype(int) - 1
        hi = np.ceil((xc.max(0) - org) / sp).astype(int) + 1
        lo = np.maximum(lo, 0)
        hi = np.minimum(hi, np.array(dims) - 1)
        if np.any(hi < lo):
            continue
        xs = org[0] + sp[0] * np.arange(lo[0], hi[0] + 1)
        ys = org[1] + sp[1] * np.arange(lo[1], hi[1] + 1)
        zs = org[2] + sp[2] * np.arange(lo[2], hi[2] + 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        P = np.stack([X, Y, Z], -1) @ A.T + t - centre
        Q = P @ rot  # coordinates in the ellipsoid frame
        inside = ((Q / axes) ** 2).sum(-1) <= 1.0
        out[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += val * inside
    return out


def remap(i):
    return (1000.0 - 0.8 * i + 0.0005 * i * i).astype(np.float32)


@dataclass
class Pair:
    fixed: Volume
    moving: Volume
    matrix: np.ndarray      # ground-truth 3x3 (moving -> fixed)
    translation: np.ndarray  # ground-truth translation (centre 0)


def make_pair(dims=(128, 128, 64), spacing=(1.5, 1.5, 1.5), n_ellipsoids=20, seed=0,
              model="rigid", noise=10.0) -> Pair:
    """Fixed/moving phantom pair: rigid as the 128x128x64 workload (rot
    U[-10,10] deg, trans U[-10,10] mm), affine as the 256x256x128 workload
    (rot +/-10 deg, scale U[0.9,1.1], shear U[-0.05,0.05], trans +/-15 mm)."""
    rng = np.random.default_rng(seed)
    origin = centred_origin(dims, spacing)
    extent = [d * s for d, s in zip(dims, spacing)]
    ells = _ellipsoids(rng, n_ellipsoids, extent_mm=extent)
    if model == "rigid":
        A, t, _ = random_rigid(rng)
    elif model == "affine":
        A, t = random_affine(rng)
    else:
        raise ValueError(model)
    fixed = _render(ells, dims, spacing, origin)
    fixed += rng.normal(0.0, noise, fixed.shape).astype(np.float32)
    moving = remap(_render(ells, dims, spacing, origin, A, t))
    moving += rng.normal(0.0, noise, moving.shape).astype(np.float32)
    return Pair(Volume(fixed, tuple(spacing), origin), Volume(moving, tuple(spacing), origin),
                A, t)


WORKLOADS = {
    # name: (dims, spacing, ellipsoids, model, block/search/pyramid settings)
    "pair128_rigid": dict(dims=(128, 128, 64), spacing=(1.5, 1.5, 1.5), n_ellipsoids=20,
                          model="rigid", nhood_width=7, nhood_gap=4, pyr_num=2),
    "pair256_affine": dict(dims=(256, 256, 128), spacing=(1.0, 1.0, 1.0), n_ellipsoids=60,
                           model="affine", nhood_width=9, nhood_gap=3, pyr_num=3),
}
