"""File I/O of the command-line drop-in (include/mirorr_io.h), CPU only.

NIfTI headers are built byte by byte here with struct (an independent
writer), with geometries whose LPS direction and origin are known in closed
form; the ARI reorientation is checked voxel by voxel for physical-position
preservation; transform files are checked against literal ITK V1.0 text and
against textbook rotations. The reference ships no image or transform files,
so there are no golden files to compare with.
"""
import gzip
import math
import struct

import numpy as np
import pytest

from mirorr_amd import capi, fileio
from mirorr_amd.phantom import Volume


def nifti_bytes(data, pixdim=(1.0, 1.0, 1.0), qform=None, sform=None, qfac=1.0,
                datatype=16, slope=0.0, inter=0.0, endian="<", ext=b"", dim0=3, extra_dims=()):
    """A NIfTI-1 single-file image. qform = (b, c, d, qx, qy, qz); sform =
    3 x 4 rows (RAS). data: numpy array (nz, ny, nx)."""
    nz, ny, nx = data.shape
    h = bytearray(348)
    struct.pack_into(endian + "i", h, 0, 348)
    dims = [dim0, nx, ny, nz, *extra_dims] + [1] * (7 - 3 - len(extra_dims))
    struct.pack_into(endian + "8h", h, 40, *dims)
    bitpix = {2: 8, 256: 8, 4: 16, 512: 16, 8: 32, 768: 32, 16: 32, 64: 64}[datatype]
    struct.pack_into(endian + "hh", h, 70, datatype, bitpix)
    struct.pack_into(endian + "8f", h, 76, qfac, *pixdim, 1, 1, 1, 1)
    vox_offset = 352 + len(ext)
    struct.pack_into(endian + "fff", h, 108, vox_offset, slope, inter)
    if qform is not None:
        struct.pack_into(endian + "h", h, 252, 1)
        struct.pack_into(endian + "6f", h, 256, *qform)
    if sform is not None:
        struct.pack_into(endian + "h", h, 254, 2)
        struct.pack_into(endian + "12f", h, 280, *np.asarray(sform, float).ravel())
    h[344:348] = b"n+1\0"
    dt = {2: "u1", 256: "i1", 4: "i2", 512: "u2", 8: "i4", 768: "u4", 16: "f4", 64: "f8"}[datatype]
    body = np.ascontiguousarray(data).astype(endian + dt).tobytes()
    return bytes(h) + b"\0\0\0\0" + ext + body


def write(tmp_path, name, blob):
    p = tmp_path / name
    if name.endswith(".gz"):
        p.write_bytes(gzip.compress(blob))
    else:
        p.write_bytes(blob)
    return str(p)


def direction(v):
    return np.asarray(v.direction).reshape(3, 3)


# ---- NIfTI reading -----------------------------------------------------------


def test_read_plain_float_identity(tmp_path):
    d = np.arange(4 * 5 * 6, dtype=np.float32).reshape(6, 5, 4)
    v = fileio.read_nifti(write(tmp_path, "a.nii", nifti_bytes(d, pixdim=(0.5, 1.5, 2.0))))
    assert v.dims == (4, 5, 6)
    assert np.allclose(v.spacing, (0.5, 1.5, 2.0))
    # No qform/sform: the voxel sizes alone; identity RAS is identity LPS
    # up to the sign of x and y.
    assert np.allclose(direction(v), np.diag([-1, -1, 1]))
    assert np.allclose(v.origin, 0)
    assert np.array_equal(v.data, d)


def test_read_qform_rotation_about_z(tmp_path):
    # Quaternion of a +90 degree rotation about z: (b, c, d) = (0, 0, sin 45).
    s = math.sin(math.pi / 4)
    d = np.zeros((2, 3, 4), np.float32)
    v = fileio.read_nifti(write(tmp_path, "q.nii.gz",
                                nifti_bytes(d, pixdim=(2, 3, 4), qform=(0, 0, s, 10, -20, 30))))
    # RAS R = [[0,-1,0],[1,0,0],[0,0,1]]; LPS = diag(-1,-1,1) R.
    assert np.allclose(direction(v), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]], atol=1e-6)
    assert np.allclose(v.origin, (-10, 20, 30))
    assert np.allclose(v.spacing, (2, 3, 4))


def test_read_qform_negative_qfac_flips_third_axis(tmp_path):
    d = np.zeros((2, 2, 2), np.float32)
    v = fileio.read_nifti(write(tmp_path, "f.nii",
                                nifti_bytes(d, qform=(0, 0, 0, 1, 2, 3), qfac=-1.0)))
    assert np.allclose(direction(v), np.diag([-1, -1, -1]))
    assert np.allclose(v.origin, (-1, -2, 3))


def test_read_sform_when_no_qform(tmp_path):
    # sform columns = direction * spacing: a sagittal-like permutation.
    sform = [[0, 0, 2.0, -5], [1.5, 0, 0, 6], [0, 1.5, 0, 7]]
    d = np.zeros((3, 3, 3), np.float32)
    v = fileio.read_nifti(write(tmp_path, "s.nii", nifti_bytes(d, pixdim=(1.5, 1.5, 2.0),
                                                               sform=sform)))
    assert np.allclose(direction(v), [[0, 0, -1], [-1, 0, 0], [0, 1, 0]])
    assert np.allclose(v.origin, (5, -6, 7))
    assert np.allclose(v.spacing, (1.5, 1.5, 2.0))


@pytest.mark.parametrize("datatype", [2, 256, 4, 512, 8, 768, 16, 64])
def test_read_datatypes_and_scaling(tmp_path, datatype):
    rng = np.random.default_rng(datatype)
    lo, hi = {2: (0, 255), 256: (-128, 127), 4: (-3000, 3000), 512: (0, 60000),
              8: (-10 ** 6, 10 ** 6), 768: (0, 10 ** 6), 16: (-5, 5), 64: (-5, 5)}[datatype]
    d = rng.integers(lo, hi, size=(3, 4, 5)).astype(np.float64)
    blob = nifti_bytes(d, datatype=datatype, slope=2.0, inter=-1.0)
    v = fileio.read_nifti(write(tmp_path, "t.nii", blob))
    assert np.allclose(v.data, (d * 2.0 - 1.0).astype(np.float32))


def test_read_big_endian_with_extension(tmp_path):
    d = np.arange(24, dtype=np.float32).reshape(2, 3, 4) - 7
    ext = bytes(range(32))
    v = fileio.read_nifti(write(tmp_path, "b.nii.gz",
                                nifti_bytes(d, endian=">", ext=ext, qform=(0, 0, 0, 1, 1, 1))))
    assert np.array_equal(v.data, d)
    assert np.allclose(v.origin, (-1, -1, 1))


def test_read_mask_casts_like_itk(tmp_path):
    d = np.array([0, 1, 2.7, 255, 256, 300, -1, 0.4], np.float32).reshape(2, 2, 2)
    m = fileio.read_mask(write(tmp_path, "m.nii", nifti_bytes(d)))
    assert m.ravel().tolist() == [0, 1, 2, 255, 0, 44, 255, 0]


def test_read_rejects_4d_and_garbage(tmp_path):
    d = np.zeros((2, 2, 2), np.float32)
    with pytest.raises(capi.MirorrError, match="3-D"):
        fileio.read_nifti(write(tmp_path, "4d.nii", nifti_bytes(d, dim0=4, extra_dims=(3,))))
    with pytest.raises(capi.MirorrError, match="NIfTI"):
        fileio.read_nifti(write(tmp_path, "junk.nii", b"x" * 400))
    blob = nifti_bytes(np.zeros((4, 4, 4), np.float32))
    with pytest.raises(capi.MirorrError, match="truncated"):
        fileio.read_nifti(write(tmp_path, "short.nii", blob[:-10]))
    with pytest.raises(capi.MirorrError, match="cannot open"):
        fileio.read_nifti(str(tmp_path / "missing.nii"))


@pytest.mark.parametrize("offset", [float("nan"), float("inf"), 3e9, 100.0])
def test_read_rejects_bad_vox_offset(tmp_path, offset):
    blob = bytearray(nifti_bytes(np.zeros((2, 2, 2), np.float32)))
    struct.pack_into("<f", blob, 108, offset)
    with pytest.raises(capi.MirorrError, match="vox_offset"):
        fileio.read_nifti(write(tmp_path, "off.nii", bytes(blob)))


@pytest.mark.parametrize("name", ["big.nii", "big.nii.gz"])
def test_read_huge_header_on_a_short_file_fails_without_allocating(tmp_path, name):
    # 32767^3 float64 voxels (280 TB) announced, 8 voxels present: the read
    # fails at the end of the file.
    blob = bytearray(nifti_bytes(np.zeros((2, 2, 2), np.float64), datatype=64))
    struct.pack_into("<3h", blob, 42, 32767, 32767, 32767)
    with pytest.raises(capi.MirorrError, match="truncated"):
        fileio.read_nifti(write(tmp_path, name, bytes(blob)))


# ---- NIfTI writing -------------------------------------------------------------


@pytest.mark.parametrize("name", ["w.nii", "w.nii.gz"])
@pytest.mark.parametrize("dirn", [
    np.eye(3),
    np.diag([-1.0, 1.0, 1.0]),                                   # det -1 (qfac = -1)
    np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1.0]]),               # ARI, det -1
    np.array([[math.cos(0.3), -math.sin(0.3), 0], [math.sin(0.3), math.cos(0.3), 0], [0, 0, 1]]),
])
def test_write_read_round_trip(tmp_path, name, dirn):
    rng = np.random.default_rng(1)
    d = rng.normal(size=(5, 6, 7)).astype(np.float32)
    v = Volume(data=d, spacing=(0.7, 1.1, 2.5), origin=(-3.5, 12.25, 40.0),
               direction=tuple(dirn.ravel()))
    p = str(tmp_path / name)
    fileio.write_nifti(p, v)
    w = fileio.read_nifti(p)
    assert w.dims == v.dims
    assert np.array_equal(w.data, d)
    assert np.allclose(w.spacing, v.spacing, atol=1e-6)
    assert np.allclose(w.origin, v.origin, atol=1e-5)
    assert np.allclose(direction(w), dirn, atol=1e-6)
    if name.endswith(".gz"):
        assert open(p, "rb").read(2) == b"\x1f\x8b"


def test_written_header_fields(tmp_path):
    v = Volume(data=np.zeros((2, 3, 4), np.float32), spacing=(1, 2, 3), origin=(4, 5, 6))
    p = str(tmp_path / "h.nii")
    fileio.write_nifti(p, v)
    h = open(p, "rb").read(352)
    assert struct.unpack_from("<i", h, 0)[0] == 348
    assert struct.unpack_from("<8h", h, 40)[:4] == (3, 4, 3, 2)
    assert struct.unpack_from("<hh", h, 70) == (16, 32)
    assert struct.unpack_from("<f", h, 108)[0] == 352.0
    assert struct.unpack_from("<hh", h, 252) == (1, 1)
    assert h[344:348] == b"n+1\0"
    # sform rows in RAS: LPS origin (4, 5, 6) is RAS (-4, -5, 6).
    srow = np.array(struct.unpack_from("<12f", h, 280)).reshape(3, 4)
    assert np.allclose(srow, [[-1, 0, 0, -4], [0, -2, 0, -5], [0, 0, 3, 6]])


# ---- ARI reorientation -----------------------------------------------------------


def _phys(dims, spacing, origin, dirn):
    nx, ny, nz = dims
    k = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    k = k.transpose(2, 1, 0, 3).reshape(-1, 3).astype(float)  # (z, y, x) order
    return origin + (k * spacing) @ np.asarray(dirn).reshape(3, 3).T


ARI = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1.0]])


@pytest.mark.parametrize("dirn", [
    np.eye(3),                                            # RAI (ITK identity)
    np.diag([-1.0, -1.0, 1.0]),                           # LPI (RAS NIfTI)
    np.array([[0, 0, -1], [-1, 0, 0], [0, 1, 0.0]]),      # sagittal slices
    np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0.0]]),       # coronal slices
    ARI,                                                  # already ARI
])
def test_reorient_ari_preserves_physical_positions(dirn):
    rng = np.random.default_rng(3)
    d = rng.normal(size=(3, 4, 5)).astype(np.float32)
    m = (rng.random((3, 4, 5)) * 255).astype(np.uint8)
    v = Volume(data=d, spacing=(0.5, 0.75, 1.25), origin=(7.0, -3.0, 2.0),
               direction=tuple(dirn.ravel()), mask=m)
    w = fileio.reorient_ari(v)
    assert np.allclose(direction(w), ARI)
    pv = _phys(v.dims, np.array(v.spacing), np.array(v.origin), v.direction)
    pw = _phys(w.dims, np.array(w.spacing), np.array(w.origin), w.direction)
    # Same multiset of (position, value): sort both by position.
    key_v = np.lexsort(np.round(pv, 6).T)
    key_w = np.lexsort(np.round(pw, 6).T)
    assert np.allclose(pv[key_v], pw[key_w])
    assert np.array_equal(d.ravel()[key_v], w.data.ravel()[key_w])
    assert np.array_equal(m.ravel()[key_v], w.mask.ravel()[key_w])


def test_reorient_ari_rejects_ambiguous_obliques():
    c = math.sqrt(0.5)
    dirn = np.array([[c, c, 0], [c, -c, 0], [0, 0, 1.0]])  # 45 degrees: ties
    v = Volume(data=np.zeros((2, 2, 2), np.float32), spacing=(1, 1, 1), origin=(0, 0, 0),
               direction=tuple(dirn.ravel()))
    with pytest.raises(capi.MirorrError, match="oblique"):
        fileio.reorient_ari(v)


# ---- centred initialisation ------------------------------------------------------


@pytest.mark.parametrize("affine", [False, True])
def test_centered_init_geometry(affine):
    mov = Volume(data=np.zeros((10, 20, 30), np.float32), spacing=(1, 2, 3), origin=(0, 0, 0))
    fix = Volume(data=np.zeros((5, 5, 5), np.float32), spacing=(2, 2, 2), origin=(100, 50, -20))
    p, c = fileio.centered_init(mov, fix, affine)
    # Geometric centres: moving (14.5, 19, 13.5), fixed (104, 54, -16).
    assert np.allclose(c, (14.5, 19.0, 13.5))
    t = p[9:] if affine else p[3:]
    assert np.allclose(t, (104 - 14.5, 54 - 19.0, -16 - 13.5))
    lin = p[:9] if affine else p[:3]
    assert np.allclose(lin, np.eye(3).ravel() if affine else 0)


# ---- transform files ---------------------------------------------------------------


def test_write_euler_literal(tmp_path):
    p = str(tmp_path / "r.tfm")
    fileio.write_tfm(p, [0.0, 0.0, 0.5, 1.0, -2.0, 3.25], [10.0, 20.0, 30.0], affine=False)
    assert open(p).read() == (
        "#Insight Transform File V1.0\n#Transform 0\n"
        "Transform: Euler3DTransform_double_3_3\n"
        "Parameters: 0 0 0.5 1 -2 3.25\nFixedParameters: 10 20 30\n")


def test_read_itk_affine_literal(tmp_path):
    p = tmp_path / "a.tfm"
    p.write_text("#Insight Transform File V1.0\n#Transform 0\n"
                 "Transform: AffineTransform_double_3_3\n"
                 "Parameters: 1.1 0.1 0 0 0.9 0 0 0 1 10 20 30\n"
                 "FixedParameters: 1 2 3\n")
    params, c = fileio.read_tfm(str(p), affine=True)
    assert np.allclose(params, [1.1, 0.1, 0, 0, 0.9, 0, 0, 0, 1, 10, 20, 30])
    assert np.allclose(c, (1, 2, 3))
    with pytest.raises(capi.AlgoError, match="non-orthogonal"):
        fileio.read_tfm(str(p), affine=False)


def test_read_euler_angles_are_textbook_rotations(tmp_path):
    # A rotation about z alone is the textbook Rz, read back as an affine matrix.
    p = tmp_path / "z.tfm"
    th = 0.3
    p.write_text("#Insight Transform File V1.0\n#Transform 0\n"
                 "Transform: Euler3DTransform_double_3_3\n"
                 f"Parameters: 0 0 {th!r} 0 0 0\nFixedParameters: 0 0 0\n")
    params, _ = fileio.read_tfm(str(p), affine=True)
    Rz = [[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]]
    assert np.allclose(np.reshape(params[:9], (3, 3)), Rz)
    # x then y: ITK's default Euler order applies Ry first, then Rx, then Rz.
    ax, ay = 0.2, -0.4
    p.write_text("#Insight Transform File V1.0\n#Transform 0\n"
                 "Transform: Euler3DTransform_double_3_3\n"
                 f"Parameters: {ax!r} {ay!r} 0 0 0 0\nFixedParameters: 0 0 0\n")
    params, _ = fileio.read_tfm(str(p), affine=True)
    Rx = np.array([[1, 0, 0], [0, math.cos(ax), -math.sin(ax)], [0, math.sin(ax), math.cos(ax)]])
    Ry = np.array([[math.cos(ay), 0, math.sin(ay)], [0, 1, 0], [-math.sin(ay), 0, math.cos(ay)]])
    assert np.allclose(np.reshape(params[:9], (3, 3)), Rx @ Ry)
    # ComputeZYX as a fourth fixed parameter: Rz Ry Rx.
    p.write_text("#Insight Transform File V1.0\n#Transform 0\n"
                 "Transform: Euler3DTransform_double_3_3\n"
                 f"Parameters: {ax!r} {ay!r} 0 0 0 0\nFixedParameters: 0 0 0 1\n")
    params, _ = fileio.read_tfm(str(p), affine=True)
    assert np.allclose(np.reshape(params[:9], (3, 3)), Ry @ Rx)


def _apply(params, c, affine, x):
    if affine:
        M = np.reshape(params[:9], (3, 3))
        t = np.asarray(params[9:])
    else:
        ax, ay, az = params[:3]
        Rx = np.array([[1, 0, 0], [0, math.cos(ax), -math.sin(ax)], [0, math.sin(ax), math.cos(ax)]])
        Ry = np.array([[math.cos(ay), 0, math.sin(ay)], [0, 1, 0], [-math.sin(ay), 0, math.cos(ay)]])
        Rz = np.array([[math.cos(az), -math.sin(az), 0], [math.sin(az), math.cos(az), 0], [0, 0, 1]])
        M = Rz @ Rx @ Ry
        t = np.asarray(params[3:])
    return (x - c) @ M.T + c + t


@pytest.mark.parametrize("affine", [False, True])
def test_tfm_round_trip_and_inverse(tmp_path, affine):
    rng = np.random.default_rng(7)
    if affine:
        ang = rng.uniform(-0.3, 0.3, 3)
        from mirorr_amd.phantom import euler_zxy
        params = np.concatenate([np.asarray(euler_zxy(*ang)).ravel(), rng.uniform(-20, 20, 3)])
    else:
        params = np.concatenate([rng.uniform(-0.3, 0.3, 3), rng.uniform(-20, 20, 3)])
    c = rng.uniform(-50, 50, 3)
    p = str(tmp_path / "x.tfm")
    fileio.write_tfm(p, params, c, affine)
    back, cb = fileio.read_tfm(p, affine)
    assert np.allclose(back, params, atol=1e-14) and np.allclose(cb, c, atol=1e-14)
    # Inverse on read and inverse on write both undo the transform (a
    # rotation, so the orthogonalisation of the inverse changes nothing).
    pts = rng.uniform(-100, 100, (20, 3))
    y = _apply(params, c, affine, pts)
    inv, ci = fileio.read_tfm(p, affine, inverse=True)
    assert np.allclose(ci, c)
    assert np.allclose(_apply(inv, ci, affine, y), pts, atol=1e-9)
    q = str(tmp_path / "inv.tfm")
    fileio.write_tfm(q, params, c, affine, inverse=True)
    inv2, ci2 = fileio.read_tfm(q, affine)
    assert np.allclose(_apply(inv2, ci2, affine, y), pts, atol=1e-9)


def test_affine_inverse_is_orthogonalised(tmp_path):
    # The reference's inverse replaces the matrix by the nearest rotation:
    # for diag(2, 2, 2) that is the identity.
    p = str(tmp_path / "s.tfm")
    fileio.write_tfm(p, [2, 0, 0, 0, 2, 0, 0, 0, 2, 4, 0, 0], [0, 0, 0], affine=True)
    inv, _ = fileio.read_tfm(p, affine=True, inverse=True)
    assert np.allclose(np.reshape(inv[:9], (3, 3)), np.eye(3))
    # translation of the exact inverse: -M^-1 t = (-2, 0, 0)
    assert np.allclose(inv[9:], (-2, 0, 0))


def test_tfm_reader_errors(tmp_path):
    p = tmp_path / "b.tfm"
    p.write_text("#Insight Transform File V1.0\n#Transform 0\n"
                 "Transform: BSplineTransform_double_3_3\nParameters: 1 2 3\n"
                 "FixedParameters: 0 0 0\n")
    with pytest.raises(capi.MirorrError, match="MatrixOffsetTransform"):
        fileio.read_tfm(str(p), affine=False)
    p.write_text("Transform: AffineTransform_double_3_3\nParameters: 1 0 0 0 1 0 0 0 x 0 0 0\n")
    with pytest.raises(capi.MirorrError, match="bad number"):
        fileio.read_tfm(str(p), affine=True)
    with pytest.raises(capi.MirorrError, match="cannot open"):
        fileio.read_tfm(str(tmp_path / "none.tfm"), affine=True)
