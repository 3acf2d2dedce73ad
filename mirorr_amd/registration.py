"""Registration with the interface of the reference's MirorrPyramidImplement
(itkMirorrPyramidImplement.h:104-139) and registration methods
(itkMirorrRegistrationMethod.h:135-276, itkSymMirorrRegistrationMethod.h),
running device-resident on the MI355X through the C ABI.

    reg = MirorrPyramid()
    reg.SetRegistrationTypeSymmetrical()   # default; or ...Classic()
    reg.SetTransformType("affine")         # "rigid" (Euler3D, default) or "affine"
    reg.SetBlockWidth(4); reg.SetNhoodWidth(9); reg.SetNhoodGap(3)
    reg.SetMaxIterations(5)                # negative halves per level
    reg.GetPyramidScheduleTuner().SetLevelMax(0)   # every level
    reg.SetFixedImage(fixed); reg.SetMovingImage(moving)
    params, center = reg.RunMirorr()        # ITK parameters + rotation centre
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .blockmatcher import Context, _live_volumes, default_context


class DeviceVolume:
    """An image (+ optional mask) resident in HBM (mirorr_volume)."""

    def __init__(self, vol, ctx: Context | None = None, mask=None, storage="f32"):
        self.ctx = ctx or default_context()
        data = np.ascontiguousarray(vol.data, np.float32)
        m = mask if mask is not None else getattr(vol, "mask", None)
        if m is not None:
            m = np.ascontiguousarray(m, np.uint8)
        g = capi.Geom.make(vol.dims, vol.spacing, vol.origin, vol.direction)
        self._h = C.c_void_p()
        st = capi.STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
        capi.check(capi.lib().mirorr_volume_create_storage(
            self.ctx.handle, capi.ptr(data), capi.ptr(m), C.byref(g), st, C.byref(self._h)),
            self.ctx.handle)
        self.dims = vol.dims
        _live_volumes.add(self)

    def info(self) -> dict:
        """Storage and, for binary16, the (offset, scale) encoding."""
        st, off, sc = C.c_int(), C.c_double(), C.c_double()
        capi.check(capi.lib().mirorr_volume_info(self._h, C.byref(st), C.byref(off),
                                                 C.byref(sc)), self.ctx.handle)
        return {"storage": st.value, "offset": off.value, "scale": sc.value}

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            capi.lib().mirorr_volume_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ScheduleTuner:
    """The level-cropping knobs of PyramidScheduleTuner (itkPyramidScheduleTuner.h:72-103)."""

    def __init__(self, params):
        self._p = params

    def SetLevelMin(self, v): self._p.pyr_start = int(v)
    def SetLevelMax(self, v): self._p.pyr_end = int(v)
    def SetMaxLevelNumber(self, v): self._p.pyr_num = int(v)
    def SetMinLength(self, v): self._p.pyr_min_size = int(v)


class MirorrPyramid:
    """MirorrPyramidImplement: pyramid + level loop + classic/symmetric method."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        p = capi.RegParams()
        capi.lib().mirorr_reg_default_params(C.byref(p))
        self._p = p
        self._fixed = self._moving = None
        self._fixed_mask = self._moving_mask = None
        self._storage = "f32"
        self.stats = None
        self._tuner = _ScheduleTuner(p)

    def SetParameters(self, params: capi.RegParams):
        """Takes a whole mirorr_reg_params (a copy)."""
        C.memmove(C.byref(self._p), C.byref(params), C.sizeof(capi.RegParams))

    # registration type / transform (itkMirorrPyramidImplement.h:104-113)
    def SetRegistrationTypeClassic(self): self._p.symmetric = 0
    def SetRegistrationTypeSymmetrical(self): self._p.symmetric = 1

    def SetTransformType(self, name: str):
        if name not in ("rigid", "affine"):
            raise ValueError("transform type must be 'rigid' or 'affine'")
        self._p.affine = 1 if name == "affine" else 0

    def SetMaxIterations(self, n: int): self._p.iterations = int(n)
    def SetBlockWidth(self, v): self._p.block_width = int(v)
    def SetBlockGap(self, v): self._p.block_gap = int(v)
    def SetNhoodWidth(self, v): self._p.nhood_width = int(v)
    def SetNhoodGap(self, v): self._p.nhood_gap = int(v)
    def SetPortionMatchesKept(self, v): self._p.portion_kept = float(v)
    def SetVarianceSelection(self, on: bool): self._p.variance_select = 1 if on else 0
    def SetVerbosity(self, v): self._p.verbosity = int(v)
    def SetCentreInitialisation(self, on: bool): self._p.centre_init = 1 if on else 0

    def SetStorage(self, name: str):
        """Device storage of the volumes, pyramid levels and half-way images:
        "f32" (the reference's float) or "f16" (binary16 with a per-volume
        affine encoding, widened to float where the search stages it)."""
        if name not in capi.STORAGE_NAMES:
            raise ValueError(f"storage must be one of {sorted(capi.STORAGE_NAMES)}")
        self._storage = name

    def SetResamplingType(self, name: str):
        """middle (default), basic, moving or fixed (SetResamplingType,
        itkMirorrPyramidImplement.txx:515-545)."""
        if name not in capi.RESAMPLING_NAMES:
            raise ValueError(f"Unknown resampling type {name}")
        self._p.resampling = capi.RESAMPLING_NAMES[name]

    def SetBlockMetricType(self, name: str):
        """normalized_correlation (default), sum_of_squared_differences or
        correlation_ratio (MirorrRegistrationMethod::SetBlockMetricType)."""
        if name not in capi.METRIC_NAMES:
            raise ValueError(f"{name} is not a valid blockmetric type")
        self._p.block_metric = capi.METRIC_NAMES[name]
    def SetStopTolerances(self, outer: float, lts: float):
        """Early-stop thresholds (1e-3 each by default); 0 disables them."""
        self._p.stop_tolerance = float(outer)
        self._p.lts_tolerance = float(lts)
    def GetPyramidScheduleTuner(self): return self._tuner

    def SetFixedImage(self, vol): self._fixed = vol
    def SetMovingImage(self, vol): self._moving = vol
    def SetFixedMask(self, m): self._fixed_mask = m
    def SetMovingMask(self, m): self._moving_mask = m

    @property
    def params(self):
        return self._p

    def RunMirorr(self, init_params=None, init_center=None, fixed_dev=None, moving_dev=None):
        """Registers moving to fixed; returns (ITK parameters, centre). Device
        volumes may be passed to skip the upload (they stay resident)."""
        fd = fixed_dev or DeviceVolume(self._fixed, self.ctx, self._fixed_mask, self._storage)
        md = moving_dev or DeviceVolume(self._moving, self.ctx, self._moving_mask, self._storage)
        n = 12 if self._p.affine else 6
        out = np.zeros(12)
        oc = np.zeros(3)
        st = capi.RegStats()
        ip = None if init_params is None else np.ascontiguousarray(init_params, np.float64)
        ic = None if init_center is None else np.ascontiguousarray(init_center, np.float64)
        capi.check(capi.lib().mirorr_register(self.ctx.handle, C.byref(self._p), fd.handle,
                                              md.handle, capi.ptr(ip), capi.ptr(ic),
                                              capi.ptr(out), capi.ptr(oc), C.byref(st)),
                   self.ctx.handle)
        self.stats = _stats_dict(st)
        return out[:n], oc


def _stats_dict(st):
    return {"levels": st.levels, "iterations": list(st.iterations[: st.levels]),
            "points": list(st.points[: st.levels]),
            "inner_iterations": st.inner_iterations, "voxel_evals": st.voxel_evals,
            "level_dims": [tuple(st.level_dims[i]) for i in range(st.levels)],
            "seconds": st.seconds, "rescored_blocks": st.rescored_blocks}


def register_batch(reg: "MirorrPyramid", fixed_devs, moving_devs, init_params=None,
                   init_centers=None):
    """mirorr_register_batch: the registrations of len(fixed_devs) pairs of
    device volumes with reg's parameters, in lockstep on reg's context.
    Returns [(params, centre)] per pair and sets reg.batch_stats."""
    n = len(fixed_devs)
    if len(moving_devs) != n:
        raise ValueError("as many moving as fixed volumes")
    fv = (C.c_void_p * max(n, 1))(*[v.handle.value for v in fixed_devs])
    mv = (C.c_void_p * max(n, 1))(*[v.handle.value for v in moving_devs])
    out = np.zeros((max(n, 1), 12))
    oc = np.zeros((max(n, 1), 3))
    st = (capi.RegStats * max(n, 1))()
    ip = ic = None
    if init_params is not None:
        ip = np.zeros((n, 12))
        for i, q in enumerate(init_params):
            ip[i, : len(q)] = q
    if init_centers is not None:
        ic = np.ascontiguousarray(init_centers, np.float64).reshape(n, 3)
    capi.check(capi.lib().mirorr_register_batch(
        reg.ctx.handle, C.byref(reg.params), n, C.cast(fv, C.c_void_p), C.cast(mv, C.c_void_p),
        capi.ptr(ip), capi.ptr(ic), capi.ptr(out), capi.ptr(oc), C.cast(st, C.c_void_p)),
        reg.ctx.handle)
    k = 12 if reg.params.affine else 6
    reg.batch_stats = [_stats_dict(st[i]) for i in range(n)]
    return [(out[i, :k].copy(), oc[i].copy()) for i in range(n)]


def lts_fit(symmetric: bool, affine: bool, X, Y, params=None, center=None, ctx=None,
            noalloc=False):
    """OptimizeTransform on given pairs (device LTS). Classic: composes into
    (params, center); symmetric: returns the half-space transforms. With
    noalloc, the variant that allocates no device memory
    (mirorr_lts_fit_noalloc, at most 1048576 pairs)."""
    ctx = ctx or default_context()
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    n = 12 if affine else 6
    p = np.zeros(12)
    if params is not None:
        p[:n] = params
    elif affine:
        p[[0, 4, 8]] = 1.0
    c = np.zeros(3) if center is None else np.ascontiguousarray(center, np.float64).copy()
    f = np.zeros((3, 4))
    b = np.zeros((3, 4))
    it = C.c_int()
    fit = capi.lib().mirorr_lts_fit_noalloc if noalloc else capi.lib().mirorr_lts_fit
    capi.check(fit(ctx.handle, int(symmetric), int(affine), capi.ptr(X), capi.ptr(Y), X.shape[0],
                   capi.ptr(p), capi.ptr(c), capi.ptr(f), capi.ptr(b), C.byref(it)), ctx.handle)
    return p[:n], c, f, b, it.value


def resample(data, gin: capi.Geom, matrix, offset, gout: capi.Geom, want_mask=False, ctx=None):
    ctx = ctx or default_context()
    data = np.ascontiguousarray(data, np.float32)
    out = np.zeros(tuple(gout.dims)[::-1], np.float32)
    m = np.zeros(tuple(gout.dims)[::-1], np.uint8) if want_mask else None
    M = np.ascontiguousarray(matrix, np.float64).reshape(9)
    off = np.ascontiguousarray(offset, np.float64).reshape(3)
    capi.check(capi.lib().mirorr_resample(ctx.handle, capi.ptr(data), C.byref(gin), capi.ptr(M),
                                          capi.ptr(off), C.byref(gout), capi.ptr(out),
                                          capi.ptr(m)), ctx.handle)
    return (out, m) if want_mask else out


_INTERP = {"nn": capi.INTERP_NN, "linear": capi.INTERP_LINEAR, "bspline": capi.INTERP_BSPLINE,
           "sinc": capi.INTERP_SINC}


def resample_final(data, gin: capi.Geom, matrix, offset, gout: capi.Geom, interpolator="bspline",
                   ctx=None):
    """Final output resampling (GetResampledImage): out(p) = data(matrix p +
    offset) on gout with the named interpolator (nn, linear, bspline, sinc)."""
    ctx = ctx or default_context()
    d = np.ascontiguousarray(data, np.float32)
    out = np.zeros(tuple(gout.dims)[::-1], np.float32)
    M = np.ascontiguousarray(matrix, np.float64).reshape(9)
    o = np.ascontiguousarray(offset, np.float64).reshape(3)
    capi.check(capi.lib().mirorr_resample_final(ctx.handle, capi.ptr(d), C.byref(gin),
                                                capi.ptr(M), capi.ptr(o), C.byref(gout),
                                                _INTERP[interpolator], capi.ptr(out)),
               ctx.handle)
    return out


def pyramid_level(data, geom: capi.Geom, schedule, level, ctx=None):
    ctx = ctx or default_context()
    data = np.ascontiguousarray(data, np.float32)
    sch = np.ascontiguousarray(schedule, np.int32).reshape(-1, 3)
    f = sch[level]
    dims = [max(1, int(np.floor(geom.dims[i] / f[i]))) for i in range(3)]
    out = np.zeros(dims[::-1], np.float32)
    og = capi.Geom()
    capi.check(capi.lib().mirorr_pyramid_level(ctx.handle, capi.ptr(data), C.byref(geom),
                                               capi.ptr(sch), sch.shape[0], int(level),
                                               capi.ptr(out), C.byref(og)), ctx.handle)
    return out, og
