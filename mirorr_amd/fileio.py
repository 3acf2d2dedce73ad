"""Image and transform files of the command-line drop-in, through the C ABI
(include/mirorr_io.h): NIfTI-1 volumes with ITK's LPS geometry, the ARI
reorientation every input gets (itkMirorrPyramidWrapper.txx:79-92, 208-211),
the centred initial transform (itkMirorrPyramidWrapper.txx:611-623) and ITK
"#Insight Transform File V1.0" files (itkMirorrPyramidWrapper.txx:353-448).

Volumes are phantom.Volume objects: float32 data of shape (nz, ny, nx).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .phantom import Volume


def _geom_tuple(g: capi.Geom):
    return (tuple(g.dims), tuple(g.spacing), tuple(g.origin), tuple(g.direction))


def _take(buf: C.c_void_p, n: int, dtype) -> np.ndarray:
    try:
        arr = np.ctypeslib.as_array(C.cast(buf, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                    shape=(n,)).copy()
    finally:
        capi.lib().mirorr_io_free(buf)
    return arr


def read_nifti(path: str, mask_path: str | None = None) -> Volume:
    """Reads a 3-D NIfTI-1 volume (.nii / .nii.gz) as float32, plus an
    optional mask volume on the same grid."""
    g = capi.Geom()
    buf = C.c_void_p()
    capi.check_io(capi.lib().mirorr_nifti_read(str(path).encode(), C.byref(g), C.byref(buf)))
    dims, sp, org, dr = _geom_tuple(g)
    data = _take(buf, dims[0] * dims[1] * dims[2], np.float32).reshape(dims[::-1])
    mask = None
    if mask_path is not None:
        mask = read_mask(mask_path)
        if mask.shape != data.shape:
            raise ValueError(f"mask {mask_path} has shape {mask.shape}, image {data.shape}")
    return Volume(data=data, spacing=sp, origin=org, direction=dr, mask=mask)


def read_mask(path: str) -> np.ndarray:
    """Reads a NIfTI-1 mask as uint8 (ITK's cast: fractions truncated,
    integers keep their low 8 bits)."""
    g = capi.Geom()
    buf = C.c_void_p()
    capi.check_io(capi.lib().mirorr_nifti_read_mask(str(path).encode(), C.byref(g),
                                                    C.byref(buf)))
    dims = tuple(g.dims)
    return _take(buf, dims[0] * dims[1] * dims[2], np.uint8).reshape(dims[::-1])


def write_nifti(path: str, vol: Volume) -> None:
    """Writes a float32 NIfTI-1 file (gzip when the name ends in .gz)."""
    data = np.ascontiguousarray(vol.data, np.float32)
    g = capi.Geom.make(vol.dims, vol.spacing, vol.origin, vol.direction)
    capi.check_io(capi.lib().mirorr_nifti_write(str(path).encode(), C.byref(g), capi.ptr(data)))


def reorient_ari(vol: Volume) -> Volume:
    """The volume (and its mask) permuted and flipped to ARI orientation,
    every voxel keeping its physical position."""
    data = np.ascontiguousarray(vol.data, np.float32).copy()
    mask = None if vol.mask is None else np.ascontiguousarray(vol.mask, np.uint8).copy()
    g = capi.Geom.make(vol.dims, vol.spacing, vol.origin, vol.direction)
    capi.check_io(capi.lib().mirorr_reorient_ari(C.byref(g), capi.ptr(data), capi.ptr(mask)))
    dims, sp, org, dr = _geom_tuple(g)
    data = data.reshape(dims[::-1])
    if mask is not None:
        mask = mask.reshape(dims[::-1])
    return Volume(data=data, spacing=sp, origin=org, direction=dr, mask=mask)


def centered_init(centre_image: Volume, other: Volume, affine: bool):
    """CenteredTransformInitializer (GeometryOn): returns (params, centre)."""
    g1 = capi.Geom.make(centre_image.dims, centre_image.spacing, centre_image.origin,
                        centre_image.direction)
    g2 = capi.Geom.make(other.dims, other.spacing, other.origin, other.direction)
    p = np.zeros(12)
    c = np.zeros(3)
    capi.check_io(capi.lib().mirorr_centered_init(C.byref(g1), C.byref(g2), int(affine),
                                                  capi.ptr(p), capi.ptr(c)))
    return p[: 12 if affine else 6], c


def write_tfm(path: str, params, center, affine: bool, inverse: bool = False) -> None:
    """Writes an ITK V1.0 transform file (Euler3D or Affine)."""
    p = np.zeros(12)
    n = 12 if affine else 6
    p[:n] = np.asarray(params, np.float64)[:n]
    c = np.ascontiguousarray(center, np.float64)
    capi.check_io(capi.lib().mirorr_tfm_write(str(path).encode(), int(affine), capi.ptr(p),
                                              capi.ptr(c), int(inverse)))


def read_tfm(path: str, affine: bool, inverse: bool = False):
    """Reads the first transform of an ITK V1.0 file as (params, centre)."""
    p = np.zeros(12)
    c = np.zeros(3)
    capi.check_io(capi.lib().mirorr_tfm_read(str(path).encode(), int(affine), capi.ptr(p),
                                             capi.ptr(c), int(inverse)))
    return p[: 12 if affine else 6], c
