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


def _render_labels(ells, dims, spacing, origin, A=None, t=None):
    """Label volume of painted ellipsoids (later ones cover earlier ones) at
    y = A x + t for every voxel centre x; label k >= 1 is the k-th ellipsoid,
    0 the background."""
    nx, ny, nz = dims
    out = np.zeros((nz, ny, nx), np.int32)
    sp = np.asarray(spacing, float)
    org = np.asarray(origin, float)
    if A is None:
        A = np.eye(3)
        t = np.zeros(3)
    Ainv = np.linalg.inv(A)
    for k, (centre, axes, rot, _val) in enumerate(ells, start=1):
        r = np.abs(rot) @ axes
        corners = np.array([[centre[0] + sx * r[0], centre[1] + sy * r[1], centre[2] + sz * r[2]]
                            for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        xc = (corners - t) @ Ainv.T
        lo = np.maximum(np.floor((xc.min(0) - org) / sp).astype(int) - 1, 0)
        hi = np.minimum(np.ceil((xc.max(0) - org) / sp).astype(int) + 1, np.array(dims) - 1)
        if np.any(hi < lo):
            continue
        xs = org[0] + sp[0] * np.arange(lo[0], hi[0] + 1)
        ys = org[1] + sp[1] * np.arange(lo[1], hi[1] + 1)
        zs = org[2] + sp[2] * np.arange(lo[2], hi[2] + 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        Q = (np.stack([X, Y, Z], -1) @ A.T + t - centre) @ rot
        inside = ((Q / axes) ** 2).sum(-1) <= 1.0
        view = out[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        view[inside] = k
    return out


def _bias_field(rng, dims, amplitude=0.2):
    """1 + amplitude * p(u) with p a random cubic polynomial in the grid's
    normalised coordinates u in [-1, 1]^3, scaled so that max |p| = 1."""
    nx, ny, nz = dims
    u = [np.linspace(-1.0, 1.0, n) for n in (nz, ny, nx)]
    Z, Y, X = np.meshgrid(*u, indexing="ij", sparse=True)
    p = np.zeros((nz, ny, nx))
    for i in range(4):
        for j in range(4 - i):
            for k in range(4 - i - j):
                p = p + rng.uniform(-1.0, 1.0) * X ** i * Y ** j * Z ** k
    p /= np.abs(p).max()
    return (1.0 + amplitude * p).astype(np.float32)


def make_ct_mr_pair(dims=(512, 512, 256), spacing=(1.0, 1.0, 1.0), n_ellipsoids=40, seed=3,
                    ct_noise=20.0, mr_noise=30.0) -> Pair:
    """CT-like fixed / MR-like moving pair under a seeded rigid transform
    (rotations U[-10,10] deg, translations U[-10,10] mm).

    Fixed: int16 Hounsfield-like values in [-1024, 3071]: air -1024 around
    ellipsoids of fat (-100), soft tissue (40) or bone (U[700, 1500]), plus
    N(0, ct_noise^2). Moving: uint16 in [0, 4095] with the class intensities
    permuted non-monotonically (air 0 < bone U[200, 500] < soft tissue 1800
    < fat 3500), a smooth cubic multiplicative bias field of +/-20% and
    Rician noise (magnitude of the signal plus complex Gaussian noise).
    Both volumes are returned as float32 with their integer values."""
    rng = np.random.default_rng(seed)
    origin = centred_origin(dims, spacing)
    extent = [d * s for d, s in zip(dims, spacing)]
    ells = _ellipsoids(rng, n_ellipsoids, axes_mm=(8.0, 70.0), extent_mm=extent)
    cls = rng.integers(1, 4, n_ellipsoids)  # 1 fat, 2 soft tissue, 3 bone
    bone = rng.uniform(700.0, 1500.0, n_ellipsoids)
    ct_val = np.where(cls == 1, -100.0, np.where(cls == 2, 40.0, bone))
    mr_val = np.where(cls == 1, 3500.0, np.where(cls == 2, 1800.0,
                                                  200.0 + 300.0 * (bone - 700.0) / 800.0))
    A, t, _ = random_rigid(rng)
    ct_lut = np.concatenate([[-1024.0], ct_val]).astype(np.float32)
    mr_lut = np.concatenate([[0.0], mr_val]).astype(np.float32)
    fixed = ct_lut[_render_labels(ells, dims, spacing, origin)]
    fixed += rng.normal(0.0, ct_noise, fixed.shape).astype(np.float32)
    fixed = np.clip(np.rint(fixed), -1024, 3071).astype(np.int16).astype(np.float32)
    moving = mr_lut[_render_labels(ells, dims, spacing, origin, A, t)] * _bias_field(rng, dims)
    n1 = rng.normal(0.0, mr_noise, moving.shape).astype(np.float32)
    n2 = rng.normal(0.0, mr_noise, moving.shape).astype(np.float32)
    moving = np.sqrt((moving + n1) ** 2 + n2 ** 2)
    moving = np.clip(np.rint(moving), 0, 4095).astype(np.uint16).astype(np.float32)
    return Pair(Volume(fixed, tuple(spacing), origin), Volume(moving, tuple(spacing), origin),
                A, t)


WORKLOADS = {
    # name: (dims, spacing, ellipsoids, model, block/search/pyramid settings)
    "pair128_rigid": dict(dims=(128, 128, 64), spacing=(1.5, 1.5, 1.5), n_ellipsoids=20,
                          model="rigid", nhood_width=7, nhood_gap=4, pyr_num=2),
    "pair256_affine": dict(dims=(256, 256, 128), spacing=(1.0, 1.0, 1.0), n_ellipsoids=60,
                           model="affine", nhood_width=9, nhood_gap=3, pyr_num=3),
    "pair512_ctmr_rigid": dict(dims=(512, 512, 256), spacing=(1.0, 1.0, 1.0), n_ellipsoids=40,
                               model="rigid", nhood_width=13, nhood_gap=3, pyr_num=4),
}
