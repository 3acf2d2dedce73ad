"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module. It loads oracle/build/libmirorr_oracle.so, which `make -C oracle`
builds from the C++ restatement of the reference algorithms in this folder.

Volumes are numpy arrays of shape (nz, ny, nx), C order, so x is the fastest
axis as in ITK's raster order. Geometry tuples are given in (x, y, z) order.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libmirorr_oracle.so")
_lib = None


class OracleGeom(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("spacing", C.c_double * 3),
                ("origin", C.c_double * 3), ("direction", C.c_double * 9)]


@dataclass
class Geom:
    """Image geometry in (x, y, z) order with a row-major direction matrix."""
    dims: tuple
    spacing: tuple = (1.0, 1.0, 1.0)
    origin: tuple = (0.0, 0.0, 0.0)
    direction: tuple = field(default_factory=lambda: tuple(np.eye(3).ravel()))

    def to_c(self) -> OracleGeom:
        g = OracleGeom()
        for i in range(3):
            g.dims[i] = int(self.dims[i])
            g.spacing[i] = float(self.spacing[i])
            g.origin[i] = float(self.origin[i])
        for i in range(9):
            g.direction[i] = float(self.direction[i])
        return g


def build() -> None:
    """Compile the oracle library (make -C oracle)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p


def _declare(L):
    L.orc_ncc_pair.restype = C.c_double
    L.orc_ncc_pair.argtypes = [_P, _P, _P, _P, _P, _P, _P, C.c_int]
    L.orc_block_grid.restype = None
    L.orc_block_grid.argtypes = [_P, C.c_int, C.c_int, _P, _P]
    L.orc_bm_get_offsets.restype = C.c_int
    L.orc_bm_get_offsets.argtypes = [_P, _P, _P, _P, C.POINTER(OracleGeom), C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P,
                                     _P, _P, _P, _P]
    L.orc_bm_candidate_count.restype = C.c_int64
    L.orc_bm_candidate_count.argtypes = [_P, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_bm_candidate_count_sel.restype = C.c_int64
    L.orc_bm_candidate_count_sel.argtypes = [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]
    L.orc_bm_get_offsets_sel.restype = C.c_int
    L.orc_bm_get_offsets_sel.argtypes = [_P, _P, _P, _P, C.POINTER(OracleGeom), C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P,
                                         _P, _P, _P, _P]
    L.orc_bm_get_offsets_metric.restype = C.c_int
    L.orc_bm_get_offsets_metric.argtypes = [_P, _P, _P, _P, C.POINTER(OracleGeom), C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int, _P,
                                            _P, _P, _P, _P, _P, _P]
    for name in ("orc_ssd_pair", "orc_cr_pair"):
        f = getattr(L, name)
        f.restype = C.c_double
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, C.c_int]
    L.orc_variance_select.restype = C.c_int64
    L.orc_variance_select.argtypes = [_P, _P, C.c_int, C.c_int, C.c_double, _P, _P]
    for name, rt, at in _EXTRA_DECLS:
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = at


class RegParams(C.Structure):
    _fields_ = [("symmetric", C.c_int), ("affine", C.c_int), ("bw", C.c_int), ("bg", C.c_int),
                ("nw", C.c_int), ("ng", C.c_int), ("iterations", C.c_int),
                ("pyr_start", C.c_int), ("pyr_end", C.c_int), ("pyr_num", C.c_int),
                ("pyr_min_size", C.c_int), ("portion_kept", C.c_double), ("nthreads", C.c_int),
                ("verbosity", C.c_int), ("centre_init", C.c_int),
                ("variance_select", C.c_int), ("stop_tolerance", C.c_double),
                ("lts_tolerance", C.c_double), ("metric", C.c_int), ("resampling", C.c_int)]


class RegStats(C.Structure):
    _fields_ = [("levels", C.c_int), ("iterations", C.c_int * 16), ("points", C.c_int64 * 16),
                ("inner_iterations", C.c_int), ("voxel_evals", C.c_int64),
                ("level_dims", (C.c_int * 3) * 16)]


_G = C.POINTER(OracleGeom)
_EXTRA_DECLS: list = [
    ("orc_resample", None, [_P, _P, _G, _P, _P, _G, _P, _P]),
    ("orc_gaussian_kernel", C.c_int, [C.c_double, C.c_double, C.c_int, _P, C.c_int]),
    ("orc_gaussian_smooth", None, [_P, _P, _P, C.c_double, _P]),
    ("orc_pyramid_level_geom", None, [_G, _P, _G]),
    ("orc_pyramid_step", None, [_P, _G, _P, _G, _P]),
    ("orc_schedule_symmetric", C.c_int, [_P, _P, _P, _P, C.c_uint, C.c_uint, _P, _P, C.c_int]),
    ("orc_cropped_level", C.c_int, [C.c_int, C.c_int, C.c_int]),
    ("orc_lts_classic", C.c_int, [C.c_int, _P, _P, C.c_int64, _P, _P]),
    ("orc_lts_symmetric", C.c_int, [C.c_int, _P, _P, C.c_int64, _P, _P, _P]),
    ("orc_sqrtm4", None, [_P, _P, C.c_double, C.c_int]),
    ("orc_orthogonalize", None, [_P, _P]),
    ("orc_rigid_quaternion", None, [_P, _P, C.c_int64, _P]),
    ("orc_quat_euler", None, [_P, _P]),
    ("orc_euler_matrix", None, [_P, _P]),
    ("orc_reg_default_params", None, [C.POINTER(RegParams)]),
    ("orc_register", C.c_int, [_P, _P, _G, _P, _P, _G, C.POINTER(RegParams), _P, _P, _P, _P,
                               C.POINTER(RegStats)]),
    ("orc_register_synthetic", C.c_int, [_G, _G, C.POINTER(RegParams), _P, _P, _P, _P, _P, _P,
                                         C.POINTER(RegStats)]),
]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(3))


def _vol(a, dtype):
    if a is None:
        return None
    a = np.asarray(a)
    if a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"expected C-contiguous {dtype} volume")
    return a


def ncc_pair(base, base_mask, search, search_mask, base_idx, search_idx, bw):
    """NCC^2 of the BW^3 blocks at voxel indices base_idx / search_idx (x, y, z)."""
    base = _vol(base, np.float32)
    search = _vol(search, np.float32)
    dims = _i3(base.shape[::-1])
    return lib().orc_ncc_pair(_ptr(base), _ptr(_vol(base_mask, np.uint8)), _ptr(search),
                              _ptr(_vol(search_mask, np.uint8)), _ptr(dims),
                              _ptr(_i3(base_idx)), _ptr(_i3(search_idx)), int(bw))


def block_grid(dims, bw, ng):
    d = _i3(dims)
    nb = np.zeros(3, np.int32)
    st = np.zeros(3, np.int32)
    lib().orc_block_grid(_ptr(d), int(bw), int(ng), _ptr(nb), _ptr(st))
    return tuple(int(v) for v in nb), tuple(int(v) for v in st)


def candidate_count(dims, bw, bg, nw, ng) -> int:
    return int(lib().orc_bm_candidate_count(_ptr(_i3(dims)), bw, bg, nw, ng))


def variance_select(base, bw=4, ng=3, portion=0.5):
    """Oracle block-variance selection of a base volume (z, y, x array).
    Returns (keep uint8 per block, keys float32 per block, number kept)."""
    base = _vol(base, np.float32)
    dims = (base.shape[2], base.shape[1], base.shape[0])
    nb, _ = block_grid(dims, bw, ng)
    n = nb[0] * nb[1] * nb[2]
    keep = np.zeros(max(n, 1), np.uint8)
    keys = np.zeros(max(n, 1), np.float32)
    k = lib().orc_variance_select(_ptr(base), _ptr(_i3(dims)), bw, ng, float(portion),
                                  _ptr(keep), _ptr(keys))
    return keep[:n], keys[:n], int(k)


def candidate_count_sel(dims, bw, bg, nw, ng, keep=None):
    kp = None if keep is None else _vol(keep, np.uint8)
    return int(lib().orc_bm_candidate_count_sel(_ptr(_i3(dims)), bw, bg, nw, ng, _ptr(kp)))


METRICS = {"nc": 0, "sd": 1, "cr": 2}


def pair_score(metric, base, base_mask, search, search_mask, base_idx, search_idx, bw):
    """One block pair under metric "nc", "sd" or "cr"."""
    if metric == "nc":
        return ncc_pair(base, base_mask, search, search_mask, base_idx, search_idx, bw)
    base = _vol(base, np.float32)
    search = _vol(search, np.float32)
    f = lib().orc_ssd_pair if metric == "sd" else lib().orc_cr_pair
    dims, bi, si = _i3(base.shape[::-1]), _i3(base_idx), _i3(search_idx)
    return f(_ptr(base), _ptr(_vol(base_mask, np.uint8)), _ptr(search),
             _ptr(_vol(search_mask, np.uint8)), _ptr(dims), _ptr(bi), _ptr(si), bw)


def bm_get_offsets(base, base_mask, search, search_mask, geom: Geom, bw=4, bg=1, nw=7,
                   ng=3, nthreads=0, keep=None, metric="nc"):
    """Oracle GetOffsets. Returns a dict with raw per-block arrays (block raster
    order) and the kept point lists in the reference's thread order. `keep`
    (uint8 per block, optional) restricts the search to the selected blocks."""
    base = _vol(base, np.float32)
    search = _vol(search, np.float32)
    if search.shape != base.shape:
        raise ValueError("base and search must share one grid")
    nb, _ = block_grid(geom.dims, bw, ng)
    n = nb[0] * nb[1] * nb[2]
    out = {
        "block_idx": np.zeros((n, 3), np.int32),
        "match_idx": np.zeros((n, 3), np.int32),
        "best": np.zeros(n, np.float32),
        "second": np.zeros(n, np.float32),
    }
    bmm = np.zeros((n, 3), np.float64)
    mmm = np.zeros((n, 3), np.float64)
    nk = C.c_int64(0)
    g = geom.to_c()
    kp = None if keep is None else _vol(keep, np.uint8)
    rc = lib().orc_bm_get_offsets_metric(_ptr(base), _ptr(_vol(base_mask, np.uint8)),
                                         _ptr(search), _ptr(_vol(search_mask, np.uint8)),
                                         C.byref(g), bw, bg, nw, ng, nthreads, _ptr(kp),
                                         METRICS[metric],
                                         _ptr(out["block_idx"]), _ptr(out["match_idx"]),
                                      _ptr(out["best"]), _ptr(out["second"]), _ptr(bmm),
                                      _ptr(mmm), C.byref(nk))
    if rc != 0:
        raise RuntimeError("oracle: no block matches occurred (empty block grid)")
    out["base_mm"] = bmm[: nk.value].copy()
    out["match_mm"] = mmm[: nk.value].copy()
    out["nblocks"] = nb
    return out


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def geom_of(vol) -> Geom:
    """Geom of a mirorr_amd.phantom.Volume-like object."""
    return Geom(dims=vol.data.shape[::-1], spacing=tuple(vol.spacing), origin=tuple(vol.origin),
                direction=tuple(vol.direction))


def resample(data, mask, gin: Geom, matrix, offset, gout: Geom, want_mask=False):
    """Linear resampling (ITK semantics) of data under y = matrix x + offset."""
    data = _vol(data, np.float32)
    out = np.zeros(tuple(gout.dims[::-1]), np.float32)
    om = np.zeros(tuple(gout.dims[::-1]), np.uint8) if want_mask else None
    a, b = gin.to_c(), gout.to_c()
    lib().orc_resample(_ptr(data), _ptr(_vol(mask, np.uint8)), C.byref(a),
                       _ptr(_f64(matrix, 9)), _ptr(_f64(offset, 3)), C.byref(b), _ptr(out),
                       _ptr(om))
    return (out, om) if want_mask else out


def gaussian_kernel(variance, max_error=0.1, max_width=32):
    buf = np.zeros(64, np.float64)
    r = lib().orc_gaussian_kernel(float(variance), float(max_error), int(max_width), _ptr(buf), 64)
    return buf[: r + 1].copy()


def gaussian_smooth(data, var, max_error=0.1):
    data = _vol(data, np.float32)
    out = np.zeros_like(data)
    lib().orc_gaussian_smooth(_ptr(data), _ptr(_i3(data.shape[::-1])), _ptr(_f64(var, 3)),
                              float(max_error), _ptr(out))
    return out


def pyramid_level_geom(g: Geom, factors) -> Geom:
    o = OracleGeom()
    a = g.to_c()
    lib().orc_pyramid_level_geom(C.byref(a), _ptr(_i3(factors)), C.byref(o))
    return Geom(dims=tuple(o.dims), spacing=tuple(o.spacing), origin=tuple(o.origin),
                direction=tuple(o.direction))


def pyramid_step(prev, gprev: Geom, ratio, gout: Geom):
    prev = _vol(prev, np.float32)
    out = np.zeros(tuple(gout.dims[::-1]), np.float32)
    a, b = gprev.to_c(), gout.to_c()
    lib().orc_pyramid_step(_ptr(prev), C.byref(a), _ptr(_i3(ratio)), C.byref(b), _ptr(out))
    return out


def schedule_symmetric(msize, mspacing, fsize, fspacing, min_length=32, max_levels=1024):
    ms = np.zeros(96, np.int32)
    fs = np.zeros(96, np.int32)
    n = lib().orc_schedule_symmetric(_ptr(_i3(msize)), _ptr(_f64(mspacing, 3)), _ptr(_i3(fsize)),
                                     _ptr(_f64(fspacing, 3)), min_length, max_levels, _ptr(ms),
                                     _ptr(fs), 32)
    return ms[: 3 * n].reshape(n, 3).tolist(), fs[: 3 * n].reshape(n, 3).tolist()


def lts_classic(kind, X, Y, params, center):
    X, Y = _f64(X), _f64(Y)
    p = _f64(params).copy()
    c = _f64(center, 3).copy()
    it = lib().orc_lts_classic(int(kind), _ptr(X), _ptr(Y), X.shape[0], _ptr(p), _ptr(c))
    return p, c, it


def lts_symmetric(kind, X, Y):
    X, Y = _f64(X), _f64(Y)
    f = np.zeros((3, 4))
    b = np.zeros((3, 4))
    p = np.zeros(12)
    it = lib().orc_lts_symmetric(int(kind), _ptr(X), _ptr(Y), X.shape[0], _ptr(f), _ptr(b), _ptr(p))
    return f, b, p[: 6 if kind == 0 else 12], it


def sqrtm4(m, eps=1e-14, max_iter=50):
    out = np.zeros((4, 4))
    lib().orc_sqrtm4(_ptr(_f64(m, (4, 4))), _ptr(out), eps, max_iter)
    return out


def orthogonalize(m):
    out = np.zeros((3, 3))
    lib().orc_orthogonalize(_ptr(_f64(m, (3, 3))), _ptr(out))
    return out


def rigid_quaternion(X, Y):
    X, Y = _f64(X), _f64(Y)
    q = np.zeros(4)
    lib().orc_rigid_quaternion(_ptr(X), _ptr(Y), X.shape[0], _ptr(q))
    return q


def quat_euler(q):
    a = np.zeros(3)
    lib().orc_quat_euler(_ptr(_f64(q, 4)), _ptr(a))
    return a


def euler_matrix(ang):
    m = np.zeros((3, 3))
    lib().orc_euler_matrix(_ptr(_f64(ang, 3)), _ptr(m))
    return m


def reg_params(**kw) -> RegParams:
    p = RegParams()
    lib().orc_reg_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def register_synthetic(fixed_geom: Geom, moving_geom: Geom, params: RegParams, matrix,
                       translation, init_params=None, init_center=None):
    """The oracle's level loop with exact correspondences from the truth
    y = matrix x + translation (moving -> fixed physical points) in place of
    the block search. Returns (params, center, stats)."""
    fg, mg = fixed_geom.to_c(), moving_geom.to_c()
    out = np.zeros(12)
    oc = np.zeros(3)
    st = RegStats()
    ip = None if init_params is None else _f64(init_params)
    ic = None if init_center is None else _f64(init_center, 3)
    rc = lib().orc_register_synthetic(C.byref(fg), C.byref(mg), C.byref(params),
                                      _ptr(_f64(matrix, 9)), _ptr(_f64(translation, 3)),
                                      _ptr(ip), _ptr(ic), _ptr(out), _ptr(oc), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle registration failed ({rc})")
    n = 12 if params.affine else 6
    stats = {"levels": st.levels, "iterations": list(st.iterations[: st.levels]),
             "points": list(st.points[: st.levels]), "inner_iterations": st.inner_iterations}
    return out[:n], oc, stats


def register(fixed, moving, params: RegParams, init_params=None, init_center=None,
             fixed_mask=None, moving_mask=None):
    """Oracle registration of two Volume objects. Returns (params, center, stats)."""
    fd = _vol(fixed.data, np.float32)
    md = _vol(moving.data, np.float32)
    fg, mg = geom_of(fixed).to_c(), geom_of(moving).to_c()
    out = np.zeros(12)
    oc = np.zeros(3)
    st = RegStats()
    ip = None if init_params is None else _f64(init_params)
    ic = None if init_center is None else _f64(init_center, 3)
    rc = lib().orc_register(_ptr(fd), _ptr(_vol(fixed_mask, np.uint8)), C.byref(fg), _ptr(md),
                            _ptr(_vol(moving_mask, np.uint8)), C.byref(mg), C.byref(params),
                            _ptr(ip), _ptr(ic), _ptr(out), _ptr(oc), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle registration failed ({rc})")
    n = 12 if params.affine else 6
    stats = {"levels": st.levels, "iterations": list(st.iterations[: st.levels]),
             "points": list(st.points[: st.levels]), "inner_iterations": st.inner_iterations,
             "voxel_evals": st.voxel_evals,
             "level_dims": [tuple(st.level_dims[i]) for i in range(st.levels)]}
    return out[:n], oc, stats
