"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed: only dlopen/dlsym)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(mirorr_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def gpu_lib():
    from mirorr_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        from mirorr_amd import build
        build.build()
    return ctypes.CDLL(capi.LIB_PATH)


def test_header_declares_entry_points():
    names = _declared()
    assert "mirorr_bm_get_offsets" in names and "mirorr_ctx_create" in names


def test_library_exports_every_declared_symbol(gpu_lib):
    missing = [n for n in sorted(_declared()) if not hasattr(gpu_lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from mirorr_amd import capi
    assert set(capi.declared_symbols()) <= _declared()
