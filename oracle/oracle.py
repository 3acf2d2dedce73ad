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
    for name, rt, at in _EXTRA_DECLS:
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = at


_EXTRA_DECLS: list = []


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


def bm_get_offsets(base, base_mask, search, search_mask, geom: Geom, bw=4, bg=1, nw=7,
                   ng=3, nthreads=0):
    """Oracle GetOffsets. Returns a dict with raw per-block arrays (block raster
    order) and the kept point lists in the reference's thread order."""
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
    rc = lib().orc_bm_get_offsets(_ptr(base), _ptr(_vol(base_mask, np.uint8)), _ptr(search),
                                  _ptr(_vol(search_mask, np.uint8)), C.byref(g), bw, bg,
                                  nw, ng, nthreads, _ptr(out["block_idx"]),
                                  _ptr(out["match_idx"]), _ptr(out["best"]),
                                  _ptr(out["second"]), _ptr(bmm), _ptr(mmm), C.byref(nk))
    if rc != 0:
        raise RuntimeError("oracle: no block matches occurred (empty block grid)")
    out["base_mm"] = bmm[: nk.value].copy()
    out["match_mm"] = mmm[: nk.value].copy()
    out["nblocks"] = nb
    return out
