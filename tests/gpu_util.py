"""Device-buffer helpers for GPU tests (through the C ABI)."""
import ctypes

import numpy as np

from compute import _capi as capi


def upload(T, array):
    a = np.ascontiguousarray(array, dtype=np.float32).ravel()
    buf = T._alloc(max(a.size, 1) * 4)
    if a.size:
        capi.check(capi.lib.cs_buffer_write(T.stream, buf.handle, 0, a.ctypes.data_as(ctypes.c_void_p), a.size * 4, 1))
    return buf


def alloc(T, n, dtype=np.float32):
    return T._alloc(max(int(n), 1) * np.dtype(dtype).itemsize)


def download(T, buf, n, dtype=np.float32):
    out = np.empty(int(n), dtype=dtype)
    if n:
        capi.check(capi.lib.cs_memcpy_dtoh(T.stream, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(buf.ptr),
                                           out.nbytes))
    return out
