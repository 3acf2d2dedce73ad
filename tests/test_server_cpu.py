"""Server entry points on a machine without a GPU: creation fails with a
platform error (no CPU fallback) and null arguments are rejected before any
device work."""
import ctypes as C

import pytest

from mirorr_amd import capi


@pytest.fixture(scope="module")
def lib():
    import os
    if not os.path.exists(capi.LIB_PATH):
        from mirorr_amd import build
        build.build()
    return capi.lib()


def test_null_arguments_are_rejected(lib):
    t = C.c_int64()
    assert lib.mirorr_submit(None, None, C.byref(t)) == capi.ERR_ARG
    assert lib.mirorr_poll(None, 1, 0.0, None) == capi.ERR_ARG
    assert lib.mirorr_server_info_get(None, None) == capi.ERR_ARG
    assert lib.mirorr_server_create(0, 0, 1, None) == capi.ERR_ARG
    h = C.c_void_p()
    assert lib.mirorr_server_create(0, 0, 65, C.byref(h)) == capi.ERR_ARG and not h
    assert b"64" in lib.mirorr_server_last_error()
    lib.mirorr_server_destroy(None)


def test_server_without_a_device_fails_loudly(lib):
    h = C.c_void_p()
    st = lib.mirorr_server_create(0, 0, 2, C.byref(h))
    if st == capi.OK:
        lib.mirorr_server_destroy(h)
        pytest.skip("a HIP device is present")
    assert st == capi.ERR_PLATFORM and not h
    assert lib.mirorr_server_last_error()
    from mirorr_amd.server import RegistrationServer
    with pytest.raises(capi.PlatformError):
        RegistrationServer(slots=2)
