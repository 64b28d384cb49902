"""The C-ABI library loads and exports every entry point include/csmi.h declares (CPU)."""
import ctypes
import os
import re

import pytest

from compute import _capi as capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "csmi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 50
    for n in names:
        assert hasattr(capi.lib, n), n
    assert set(names) == set(capi.EXPORTED)


def test_version_and_errors_without_device():
    assert b"gfx950" in capi.lib.cs_version()
    n = ctypes.c_int(-1)
    assert capi.lib.cs_device_count(ctypes.byref(n)) == 0
    if n.value == 0:
        dev = ctypes.c_void_p()
        st = capi.lib.cs_device_open(0, ctypes.byref(dev))
        assert st == -1  # CL_DEVICE_NOT_FOUND
        with pytest.raises(capi.DeviceNotFound):
            capi.check(st, "open")
        assert "ordinal" in capi.last_error()


def test_status_names():
    assert capi.lib.cs_status_string(0) == b"CS_SUCCESS"
    assert capi.lib.cs_status_string(-11) == b"CS_BUILD_PROGRAM_FAILURE"
    assert capi.lib.cs_status_string(-54) == b"CS_INVALID_WORK_GROUP_SIZE"
    assert capi.lib.cs_status_string(-2) == b"CS_UNKNOWN_STATUS"
    for status, cls in capi._BY_STATUS.items():  # every mapped code has its own name
        assert capi.lib.cs_status_string(status) != b"CS_UNKNOWN_STATUS", cls


def test_null_arguments_are_rejected():
    assert capi.lib.cs_stream_synchronize(None) == -30
    assert capi.lib.cs_buffer_release(None) == -38
    assert capi.lib.cs_event_release(None) == -58
    assert capi.lib.cs_fused_release(None) == -30
    with pytest.raises(capi.InvalidValue):
        capi.check(capi.lib.cs_reduce_sum(None, None, 0, None, None))


def test_program_check_reports_build_log():
    log = ctypes.c_void_p()
    st = capi.lib.cs_program_check(b"extern \"C\" __global__ void k(float* o) { o[0] = undefined_name; }",
                                   None, ctypes.byref(log))
    assert st == -11
    assert "undefined_name" in capi.take_string(log.value)
    st = capi.lib.cs_program_check(b"extern \"C\" __global__ void k(float* o) { o[threadIdx.x] = 1.0f; }",
                                   None, ctypes.byref(log))
    assert st == 0
