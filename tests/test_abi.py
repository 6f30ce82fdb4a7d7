"""C-ABI library: loads, exports every declared symbol, validates arguments
before touching a device (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vx", "vx_api.h")).read()
    return sorted(set(re.findall(r"^\s*(?:vx_status|int32_t|const char\*)\s+(vx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from voxelengine_amd import _lib
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)
    assert L.vx_api_version() == 1


def test_null_arguments_are_rejected():
    from voxelengine_amd import _lib
    L = _lib.lib()
    out = C.c_void_p()
    assert L.vx_ctx_create(0, None) == _lib.VX_ERR_INVALID_ARGUMENT
    assert b"NULL" in L.vx_last_error()
    assert L.vx_model_create(None, 4, 4, 4, C.byref(out)) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_volume_create(None, 4, 4, 4, 2, C.byref(out)) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_trace(None, None, None, None, None, None, None, None) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_trace_rays(None, None, None, 0, 0, None, None) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_primary(None, None, None, None, None) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_volume_rebuild(None, None, 0, 16, None, None, None) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_shade_forward(None, None, None, None, None, None, None, None, None) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_ctx_destroy(None) == 0
    assert L.vx_volume_destroy(None) == 0
    assert L.vx_model_destroy(None) == 0


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from voxelengine_amd import _lib
    L = _lib.lib()
    out = C.c_void_p()
    assert L.vx_ctx_create(0, C.byref(out)) == _lib.VX_ERR_DEVICE
    assert not out.value


def test_failure_injection_arguments():
    """vx_debug_inject_failure checks its kind and count; -1 disarms."""
    from voxelengine_amd import _lib
    L = _lib.lib()
    assert L.vx_debug_inject_failure(7, 0) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_debug_inject_failure(0, -2) == _lib.VX_ERR_INVALID_ARGUMENT
    for kind in (0, 1):
        assert L.vx_debug_inject_failure(kind, 3) == 0
        assert L.vx_debug_inject_failure(kind, -1) == 0
    out = C.c_void_p()
    assert L.vx_batch_update(None, None, 0) == _lib.VX_ERR_INVALID_ARGUMENT
    assert L.vx_batch_info(None, None, None) == _lib.VX_ERR_INVALID_ARGUMENT
