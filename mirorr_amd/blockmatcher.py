"""Block matcher with the interface of the reference's itk::AbstractBlockMatcher
(itkAbstractBlockMatcher.h:41-313), running on the MI355X through the C ABI.

    bm = BlockMatcher()
    bm.SetBlockWidth(4); bm.SetNhoodWidth(7); bm.SetNhoodGap(3)
    bm.SetBaseImage(moving); bm.SetSearchImage(fixed)
    base_pts, match_pts = bm.GetOffsets()    # (n, 3) mm, like PointListType
    bm.SwapImageOrder()                      # no re-upload

Images are mirorr_amd.phantom.Volume objects (float32 (nz, ny, nx) data with
spacing/origin/direction in (x, y, z) order); masks are uint8 arrays of the
same shape or None (= all 255).
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import capi

# Handles are released before interpreter shutdown (the HIP runtime may be
# torn down before Python's own finalisation reaches __del__).
_live_matchers: "weakref.WeakSet" = weakref.WeakSet()
_live_volumes: "weakref.WeakSet" = weakref.WeakSet()
_live_contexts: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_all():
    for m in list(_live_matchers):
        m.close()
    for v in list(_live_volumes):
        v.close()
    for c in list(_live_contexts):
        c.close()


class Context:
    """A device context (HIP stream + HBM budget) shared by handles."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = C.c_void_p()
        capi.check(capi.lib().mirorr_ctx_create(device, hbm_budget, C.byref(self._h)))
        _live_contexts.add(self)

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        capi.check(capi.lib().mirorr_ctx_synchronize(self._h), self._h)

    PROF_FAST = 1    # MIRORR_PROF_FAST: the all-masks-on search kernel
    PROF_MASKED = 2  # MIRORR_PROF_MASKED: the masked search kernel
    PROF_FINEST_ONLY = 4  # registrations: instrument the full-resolution level only

    def set_profiling(self, on):
        """True instruments both search kernels, False none; an int selects
        them by bit (PROF_FAST | PROF_MASKED)."""
        mask = (self.PROF_FAST | self.PROF_MASKED) if on is True else int(on)
        capi.check(capi.lib().mirorr_ctx_set_profiling(self._h, mask), self._h)

    @property
    def iteration_graphs(self) -> bool:
        """Whether registrations replay each level's later outer iterations
        from a captured HIP graph (results are identical either way)."""
        return bool(capi.lib().mirorr_ctx_iteration_graphs(self._h))

    @iteration_graphs.setter
    def iteration_graphs(self, on: bool):
        capi.check(capi.lib().mirorr_ctx_set_iteration_graphs(self._h, 1 if on else 0), self._h)

    def profile(self, kernel: str):
        """(launches, total ms, work units) of an instrumented kernel."""
        n = C.c_int64()
        ms = C.c_double()
        w = C.c_double()
        capi.check(capi.lib().mirorr_ctx_profile_get(self._h, kernel.encode(), C.byref(n),
                                                     C.byref(ms), C.byref(w)), self._h)
        return int(n.value), float(ms.value), float(w.value)

    def profile_reset(self):
        capi.check(capi.lib().mirorr_ctx_profile_reset(self._h), self._h)

    def close(self):
        if self._h:
            capi.lib().mirorr_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _geom_of(vol):
    return capi.Geom.make(vol.dims, vol.spacing, vol.origin, vol.direction)


class BlockMatcher:
    """GPU block matcher (the BlockMatcherGPU role of the reference)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        L = capi.lib()
        p = capi.BMParams()
        L.mirorr_bm_default_params(C.byref(p))
        self._p = p
        self._h = C.c_void_p()
        capi.check(L.mirorr_bm_create(self.ctx.handle, C.byref(p), C.byref(self._h)),
                   self.ctx.handle)
        self._img = [None, None]
        self._mask = [None, None]
        self._dirty = [True, True]
        _live_matchers.add(self)

    # -- parameters (itkAbstractBlockMatcher.h:101-127) --------------------
    def _set(self, name, value):
        setattr(self._p, name, value)
        capi.check(capi.lib().mirorr_bm_set_params(self._h, C.byref(self._p)), self.ctx.handle)

    def SetBlockWidth(self, v): self._set("block_width", int(v))
    def GetBlockWidth(self): return self._p.block_width
    def SetBlockGap(self, v): self._set("block_gap", int(v))
    def GetBlockGap(self): return self._p.block_gap
    def SetNhoodWidth(self, v): self._set("nhood_width", int(v))
    def GetNhoodWidth(self): return self._p.nhood_width
    def SetNhoodGap(self, v): self._set("nhood_gap", int(v))
    def GetNhoodGap(self): return self._p.nhood_gap
    def SetVerbosity(self, v): self._set("verbosity", int(v))
    def GetVerbosity(self): return self._p.verbosity

    def SetPortionMatchesKept(self, v):
        # itkAbstractBlockMatcher.txx:83-99: 0 means 0.5, outside [0,1] raises.
        v = float(v)
        if v == 0.0:
            v = 0.5
        if v < 0.0 or v > 1.0:
            raise ValueError(f"itkBlockMatcher::SetPortionMatchesKept input is {v}. "
                             "It should be between 0 and 1.0.")
        self._set("portion_kept", v)

    def GetPortionMatchesKept(self): return self._p.portion_kept
    def SetVarianceSelection(self, on: bool): self._set("variance_select", 1 if on else 0)

    # Exact double re-scoring of the blocks next to the score filter's
    # thresholds (on by default); off leaves the float scores as searched.
    def SetExactRescore(self, on: bool): self._set("exact_rescore", 1 if on else 0)

    # Device storage of images set afterwards: "f32" (default) or "f16"
    # (binary16 with a per-image affine encoding; NCC, 4-voxel blocks).
    def SetStorage(self, name: str):
        self._set("storage", capi.STORAGE_NAMES[name])
        self._dirty = [True, True]  # re-uploaded in the new storage

    def SetBlockMetricType(self, name: str):
        """normalized_correlation (default), sum_of_squared_differences or
        correlation_ratio (BlockMatcher::SetBlockMetricType, itkBlockMatcher.h:143-175)."""
        if name not in capi.METRIC_NAMES:
            raise ValueError(f"{name} is not a valid blockmetric type")
        self._set("block_metric", capi.METRIC_NAMES[name])

    def GetBlockMetric(self): return self._p.block_metric

    # -- images (itkAbstractBlockMatcher.h:85-95) --------------------------
    def SetBaseImage(self, vol):
        self._img[0] = vol
        self._dirty[0] = True

    def SetSearchImage(self, vol):
        self._img[1] = vol
        self._dirty[1] = True

    def SetBaseMask(self, mask):
        self._mask[0] = mask
        self._dirty[0] = True

    def SetSearchMask(self, mask):
        self._mask[1] = mask
        self._dirty[1] = True

    def GetBaseImage(self): return self._img[0]
    def GetSearchImage(self): return self._img[1]

    def SwapImageOrder(self):
        """itkAbstractBlockMatcher.txx:101-112; device data is swapped, not copied."""
        self._upload()
        self._img.reverse()
        self._mask.reverse()
        capi.check(capi.lib().mirorr_bm_swap(self._h), self.ctx.handle)

    def _upload(self):
        L = capi.lib()
        for role in (capi.ROLE_BASE, capi.ROLE_SEARCH):
            if not self._dirty[role]:
                continue
            vol = self._img[role]
            if vol is None:
                raise RuntimeError("base and search images must be set")
            data = np.ascontiguousarray(vol.data, np.float32)
            mask = self._mask[role]
            if mask is None:
                mask = getattr(vol, "mask", None)
            if mask is not None:
                mask = np.ascontiguousarray(mask, np.uint8)
                if mask.shape != data.shape:
                    raise ValueError("mask and image shapes differ")
            g = _geom_of(vol)
            capi.check(L.mirorr_bm_set_volume(self._h, role, capi.ptr(data), capi.ptr(mask),
                                              C.byref(g)), self.ctx.handle)
            self._dirty[role] = False

    # -- matching ------------------------------------------------------------
    def block_count(self):
        self._upload()
        n = C.c_size_t()
        grid = (C.c_int * 3)()
        capi.check(capi.lib().mirorr_bm_block_count(self._h, C.byref(n), grid), self.ctx.handle)
        return int(n.value), tuple(grid)

    def GetOffsets(self, with_scores: bool = False):
        """Kept block centres (base) and matched centres (search), (n, 3) mm."""
        n, _ = self.block_count()
        bmm = np.zeros((n, 3), np.float64)
        mmm = np.zeros((n, 3), np.float64)
        best = np.zeros(n, np.float32)
        second = np.zeros(n, np.float32)
        index = np.zeros(n, np.int32)
        k = C.c_size_t()
        capi.check(capi.lib().mirorr_bm_get_offsets(self._h, capi.ptr(bmm), capi.ptr(mmm), n,
                                                    C.byref(k), capi.ptr(best),
                                                    capi.ptr(second), capi.ptr(index)),
                   self.ctx.handle)
        k = k.value
        if with_scores:
            return bmm[:k], mmm[:k], {"best": best[:k], "second": second[:k],
                                      "block_index": index[:k]}
        return bmm[:k], mmm[:k]

    def get_raw(self):
        """Unfiltered per-block (block_idx, match_idx, best, second), raster order."""
        n, grid = self.block_count()
        bi = np.zeros((n, 3), np.int32)
        mi = np.zeros((n, 3), np.int32)
        best = np.zeros(n, np.float32)
        second = np.zeros(n, np.float32)
        nn = C.c_size_t()
        capi.check(capi.lib().mirorr_bm_get_raw(self._h, capi.ptr(bi), capi.ptr(mi),
                                                capi.ptr(best), capi.ptr(second), n,
                                                C.byref(nn)), self.ctx.handle)
        return {"block_idx": bi, "match_idx": mi, "best": best, "second": second,
                "nblocks": grid}

    def match_async(self):
        """Enqueue one full search (statistics, preparation, NCC) without waiting."""
        self._upload()
        capi.check(capi.lib().mirorr_bm_match_async(self._h), self.ctx.handle)

    def voxel_evals(self) -> int:
        self._upload()
        v = C.c_int64()
        capi.check(capi.lib().mirorr_bm_voxel_evals(self._h, C.byref(v)), self.ctx.handle)
        return int(v.value)

    def close(self):
        if self._h:
            capi.lib().mirorr_bm_destroy(self._h)
            self._h = C.c_void_p()
        self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
