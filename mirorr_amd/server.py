"""Registration server on one GPU (mirorr_server_* / mirorr_submit /
mirorr_poll in include/mirorr_gpu.h).

    with RegistrationServer(slots=4) as srv:
        t = srv.submit(fixed, moving, params)      # params: a MirorrPyramid's
        r = srv.poll(t)                            # waits; dict with params, center, ...

Requests are admitted onto worker slots whenever a slot is free and the
request's device-memory bound fits in the server's HBM budget, in arrival
order except that later requests that fit may overtake one waiting for
memory (at most 8 times); admitted requests run at the same time, each on
its own HIP stream. There is
no counterpart in the reference, whose command line runs one pair per
process: this is the loop a service puts in front of MirorrPyramid.RunMirorr.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


def _result_dict(r: capi.RequestResult, affine: bool) -> dict:
    st = r.stats
    n = 12 if affine else 6
    return {
        "status": r.status,
        "error": r.error.decode(errors="replace"),
        "params": np.array(r.params[:n]),
        "center": np.array(r.center[:3]),
        "t_submit": r.t_submit, "t_start": r.t_start, "t_end": r.t_end,
        "latency": r.t_end - r.t_submit,
        "slot": r.slot, "bytes_bound": r.bytes_bound,
        "stats": {"levels": st.levels, "iterations": list(st.iterations[: st.levels]),
                  "points": list(st.points[: st.levels]),
                  "inner_iterations": st.inner_iterations, "voxel_evals": st.voxel_evals,
                  "seconds": st.seconds, "rescored_blocks": st.rescored_blocks},
    }


class RegistrationServer:
    """Owns a mirorr_server: `slots` workers sharing an HBM budget of
    `hbm_budget` bytes (0: the device's free memory less a reserve)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, slots: int = 4, batch: int = 1):
        """`batch` > 1: a slot runs up to that many pending requests with the
        same parameters together (mirorr_server_set_batch)."""
        self._h = C.c_void_p()
        capi.check_server(capi.lib().mirorr_server_create(int(device), int(hbm_budget),
                                                          int(slots), C.byref(self._h)))
        if batch != 1:
            capi.check_server(capi.lib().mirorr_server_set_batch(self._h, int(batch)))
        self._affine = {}
        self._borrowed = {}  # ticket -> arrays the library reads until the result is polled

    def submit(self, fixed, moving, params: capi.RegParams, fixed_mask=None, moving_mask=None,
               init_params=None, init_center=None, borrow: bool = False,
               storage: str = "f32") -> int:
        """Queues one registration of `moving` to `fixed` (Volume-like objects
        with data/dims/spacing/origin/direction); returns its ticket. The
        library copies the arrays, or with `borrow` uploads from them when the
        request runs: this object then keeps them referenced until poll()
        returns the result, and the caller must not change them meanwhile.
        `storage` "f16" keeps both volumes in binary16 (see
        MirorrPyramid.SetStorage)."""
        q = capi.Request()
        q.params = params
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return capi.ptr(a)

        fm = fixed_mask if fixed_mask is not None else getattr(fixed, "mask", None)
        mm = moving_mask if moving_mask is not None else getattr(moving, "mask", None)
        q.fixed = arr(fixed.data, np.float32)
        q.fixed_mask = arr(fm, np.uint8)
        q.fixed_geom = capi.Geom.make(fixed.dims, fixed.spacing, fixed.origin, fixed.direction)
        q.moving = arr(moving.data, np.float32)
        q.moving_mask = arr(mm, np.uint8)
        q.moving_geom = capi.Geom.make(moving.dims, moving.spacing, moving.origin,
                                       moving.direction)
        q.init_params = arr(init_params, np.float64)
        q.init_center = arr(init_center, np.float64)
        q.borrow = 1 if borrow else 0
        q.storage = capi.STORAGE_NAMES[storage]
        t = C.c_int64()
        capi.check_server(capi.lib().mirorr_submit(self._h, C.byref(q), C.byref(t)))
        self._affine[t.value] = bool(params.affine)
        if borrow:
            self._borrowed[t.value] = keep
        return t.value

    def poll(self, ticket: int, timeout: float | None = None):
        """The ticket's result dict, or None if it has not finished within
        `timeout` seconds (None: wait for it; 0: do not wait). A failed
        registration is returned with its status and message, not raised."""
        r = capi.RequestResult()
        t = -1.0 if timeout is None else float(timeout)
        st = capi.lib().mirorr_poll(self._h, int(ticket), t, C.byref(r))
        if st == capi.PENDING:
            return None
        capi.check_server(st)
        self._borrowed.pop(int(ticket), None)
        return _result_dict(r, self._affine.pop(int(ticket)))

    def info(self) -> dict:
        i = capi.ServerInfo()
        capi.check_server(capi.lib().mirorr_server_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in capi.ServerInfo._fields_}

    def close(self):
        if self._h:
            capi.lib().mirorr_server_destroy(self._h)  # waits for the running requests
            self._h = C.c_void_p()
            self._borrowed.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
