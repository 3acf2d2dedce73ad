"""ctypes binding of the C ABI (include/mirorr_gpu.h, include/mirorr_io.h).

Loads mirorr_amd/lib/libmirorr_gpu.so, built by mirorr_amd/build.py. There is
no fallback: if the library is missing, importing the device path raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmirorr_gpu.so")

OK, ERR_ARG, ERR_PLATFORM, ERR_ALGO, ERR_NOMEM, ERR_STATE, PENDING = range(7)
ROLE_BASE, ROLE_SEARCH = 0, 1
INTERP_NN, INTERP_LINEAR, INTERP_BSPLINE, INTERP_SINC = range(4)
METRIC_NC, METRIC_SSD, METRIC_CR = range(3)
RESAMPLE_MIDDLE, RESAMPLE_BASIC, RESAMPLE_MOVING, RESAMPLE_FIXED = range(4)
STORAGE_F32, STORAGE_F16 = 0, 1
STORAGE_NAMES = {"f32": STORAGE_F32, "float32": STORAGE_F32, "f16": STORAGE_F16,
                 "float16": STORAGE_F16}
RESAMPLING_NAMES = {"middle": RESAMPLE_MIDDLE, "basic": RESAMPLE_BASIC,
                    "moving": RESAMPLE_MOVING, "fixed": RESAMPLE_FIXED}
# BlockMatcher's metric names (itkBlockMatcher.h:316-323) and the CLI's short ones.
METRIC_NAMES = {"normalized_correlation": METRIC_NC, "nc": METRIC_NC,
                "sum_of_squared_differences": METRIC_SSD, "sd": METRIC_SSD,
                "correlation_ratio": METRIC_CR, "cr": METRIC_CR}


class MirorrError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class PlatformError(MirorrError):
    """Device failure (the reference's BlockMatcherGPUExceptionPlatform)."""


class AlgoError(MirorrError):
    """No block matches (the reference's BlockMatcherGPUExceptionAlgo)."""


class Geom(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("spacing", C.c_double * 3),
                ("origin", C.c_double * 3), ("direction", C.c_double * 9)]

    @classmethod
    def make(cls, dims, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), direction=None):
        g = cls()
        for i in range(3):
            g.dims[i] = int(dims[i])
            g.spacing[i] = float(spacing[i])
            g.origin[i] = float(origin[i])
        d = direction if direction is not None else (1, 0, 0, 0, 1, 0, 0, 0, 1)
        for i in range(9):
            g.direction[i] = float(d[i])
        return g


class BMParams(C.Structure):
    _fields_ = [("block_width", C.c_int), ("block_gap", C.c_int), ("nhood_width", C.c_int),
                ("nhood_gap", C.c_int), ("portion_kept", C.c_double),
                ("variance_select", C.c_int), ("verbosity", C.c_int),
                ("block_metric", C.c_int), ("exact_rescore", C.c_int), ("storage", C.c_int)]


class RegParams(C.Structure):
    _fields_ = [("symmetric", C.c_int), ("affine", C.c_int), ("block_width", C.c_int),
                ("block_gap", C.c_int), ("nhood_width", C.c_int), ("nhood_gap", C.c_int),
                ("iterations", C.c_int), ("pyr_start", C.c_int), ("pyr_end", C.c_int),
                ("pyr_num", C.c_int), ("pyr_min_size", C.c_int), ("portion_kept", C.c_double),
                ("variance_select", C.c_int), ("verbosity", C.c_int), ("centre_init", C.c_int),
                ("stop_tolerance", C.c_double), ("lts_tolerance", C.c_double),
                ("block_metric", C.c_int), ("resampling", C.c_int)]


class RegStats(C.Structure):
    _fields_ = [("levels", C.c_int), ("iterations", C.c_int * 16), ("points", C.c_int64 * 16),
                ("inner_iterations", C.c_int), ("voxel_evals", C.c_int64),
                ("level_dims", (C.c_int * 3) * 16), ("seconds", C.c_double),
                ("rescored_blocks", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("params", RegParams), ("fixed", C.c_void_p), ("fixed_mask", C.c_void_p),
                ("fixed_geom", Geom), ("moving", C.c_void_p), ("moving_mask", C.c_void_p),
                ("moving_geom", Geom), ("init_params", C.c_void_p), ("init_center", C.c_void_p),
                ("borrow", C.c_int), ("storage", C.c_int)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("params", C.c_double * 12), ("center", C.c_double * 3),
                ("stats", RegStats), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double), ("slot", C.c_int), ("bytes_bound", C.c_size_t),
                ("error", C.c_char * 256)]


class ServerInfo(C.Structure):
    _fields_ = [("slots", C.c_int), ("pending", C.c_int), ("running", C.c_int),
                ("completed", C.c_int64), ("hbm_budget", C.c_size_t), ("hbm_used", C.c_size_t),
                ("hbm_peak", C.c_size_t), ("max_running", C.c_int),
                ("reserve_peak", C.c_size_t), ("batch_max", C.c_int), ("max_batch", C.c_int)]


_P = C.c_void_p
_lib = None

# name: (restype, argtypes)
_SIGS = {
    "mirorr_ctx_create": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(_P)]),
    "mirorr_ctx_destroy": (None, [_P]),
    "mirorr_last_error": (C.c_char_p, [_P]),
    "mirorr_ctx_synchronize": (C.c_int, [_P]),
    "mirorr_ctx_stream": (_P, [_P]),
    "mirorr_ctx_set_profiling": (C.c_int, [_P, C.c_int]),
    "mirorr_ctx_set_iteration_graphs": (C.c_int, [_P, C.c_int]),
    "mirorr_ctx_iteration_graphs": (C.c_int, [_P]),
    "mirorr_ctx_profile_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mirorr_ctx_profile_reset": (C.c_int, [_P]),
    "mirorr_bm_default_params": (None, [C.POINTER(BMParams)]),
    "mirorr_bm_create": (C.c_int, [_P, C.POINTER(BMParams), C.POINTER(_P)]),
    "mirorr_bm_destroy": (None, [_P]),
    "mirorr_bm_set_params": (C.c_int, [_P, C.POINTER(BMParams)]),
    "mirorr_bm_set_volume": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(Geom)]),
    "mirorr_bm_set_volume_device": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(Geom)]),
    "mirorr_bm_swap": (C.c_int, [_P]),
    "mirorr_bm_get_offsets": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t), _P, _P,
                                        _P]),
    "mirorr_bm_block_count": (C.c_int, [_P, C.POINTER(C.c_size_t), _P]),
    "mirorr_bm_get_raw": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "mirorr_bm_match_async": (C.c_int, [_P]),
    "mirorr_bm_voxel_evals": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "mirorr_reg_default_params": (None, [C.POINTER(RegParams)]),
    "mirorr_volume_create": (C.c_int, [_P, _P, _P, C.POINTER(Geom), C.POINTER(_P)]),
    "mirorr_volume_create_storage": (C.c_int, [_P, _P, _P, C.POINTER(Geom), C.c_int,
                                               C.POINTER(_P)]),
    "mirorr_volume_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "mirorr_half_encode": (C.c_int, [_P, C.c_size_t, _P, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "mirorr_tile_force": (C.c_int, [C.c_int, C.c_int]),
    "mirorr_tile_autotune_report": (C.c_int, [C.c_char_p, C.c_size_t]),
    "mirorr_volume_destroy": (None, [_P]),
    "mirorr_register": (C.c_int, [_P, C.POINTER(RegParams), _P, _P, _P, _P, _P, _P,
                                  C.POINTER(RegStats)]),
    "mirorr_register_batch": (C.c_int, [_P, C.POINTER(RegParams), C.c_int, _P, _P, _P, _P, _P,
                                        _P, _P]),
    "mirorr_lts_fit": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, C.c_size_t, _P, _P, _P, _P,
                                 C.POINTER(C.c_int)]),
    "mirorr_lts_fit_noalloc": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, C.c_size_t, _P, _P, _P,
                                         _P, C.POINTER(C.c_int)]),
    "mirorr_resample": (C.c_int, [_P, _P, C.POINTER(Geom), _P, _P, C.POINTER(Geom), _P, _P]),
    "mirorr_pyramid_level": (C.c_int, [_P, _P, C.POINTER(Geom), _P, C.c_int, C.c_int, _P,
                                       C.POINTER(Geom)]),
    "mirorr_resample_final": (C.c_int, [_P, _P, C.POINTER(Geom), _P, _P, C.POINTER(Geom),
                                        C.c_int, _P]),
    "mirorr_server_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(_P)]),
    "mirorr_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "mirorr_poll": (C.c_int, [_P, C.c_int64, C.c_double, C.POINTER(RequestResult)]),
    "mirorr_server_info_get": (C.c_int, [_P, C.POINTER(ServerInfo)]),
    "mirorr_server_set_batch": (C.c_int, [_P, C.c_int]),
    "mirorr_server_last_error": (C.c_char_p, []),
    "mirorr_server_destroy": (None, [_P]),
    # include/mirorr_io.h
    "mirorr_io_last_error": (C.c_char_p, []),
    "mirorr_io_free": (None, [_P]),
    "mirorr_nifti_read": (C.c_int, [C.c_char_p, C.POINTER(Geom), C.POINTER(_P)]),
    "mirorr_nifti_read_mask": (C.c_int, [C.c_char_p, C.POINTER(Geom), C.POINTER(_P)]),
    "mirorr_nifti_write": (C.c_int, [C.c_char_p, C.POINTER(Geom), _P]),
    "mirorr_reorient_ari": (C.c_int, [C.POINTER(Geom), _P, _P]),
    "mirorr_centered_init": (C.c_int, [C.POINTER(Geom), C.POINTER(Geom), C.c_int, _P, _P]),
    "mirorr_tfm_write": (C.c_int, [C.c_char_p, C.c_int, _P, _P, C.c_int]),
    "mirorr_tfm_read": (C.c_int, [C.c_char_p, C.c_int, _P, _P, C.c_int]),
}


def declared_symbols():
    return sorted(_SIGS)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library not built: {LIB_PATH}. Run `python -m mirorr_amd.build` "
                "(there is no CPU fallback).")
        L = C.CDLL(LIB_PATH)
        for name, (rt, at) in _SIGS.items():
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = at
        _lib = L
    return _lib


def check(status: int, ctx=None):
    if status == OK:
        return
    msg = lib().mirorr_last_error(ctx).decode() if ctx else "error"
    cls = {ERR_PLATFORM: PlatformError, ERR_ALGO: AlgoError}.get(status, MirorrError)
    raise cls(status, msg)


def check_server(status: int):
    """Raises for a failed server call (mirorr_server_create / _submit / _poll)."""
    if status == OK:
        return
    msg = lib().mirorr_server_last_error().decode()
    cls = {ERR_PLATFORM: PlatformError, ERR_ALGO: AlgoError}.get(status, MirorrError)
    raise cls(status, msg)


def check_io(status: int):
    """Raises for a failed call of the file I/O part (include/mirorr_io.h)."""
    if status == OK:
        return
    msg = lib().mirorr_io_last_error().decode()
    cls = {ERR_PLATFORM: PlatformError, ERR_ALGO: AlgoError}.get(status, MirorrError)
    raise cls(status, msg)


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def tile_autotune_report() -> str:
    """The NCC search's tile-shape choices so far (one line per key)."""
    n = lib().mirorr_tile_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().mirorr_tile_autotune_report(buf, n + 1)
    return buf.value.decode()
