"""cs_reduce_rows_noalloc, the row reduction that allocates nothing (the FP64 partials of its
total sit in a fixed device array of the library), against cs_reduce_rows: the same bits
for row sums, row maxima and both totals, on two calls in a row, the second with a smaller
input; rows beyond its limit are refused when a total is asked for."""
import ctypes

import numpy as np
import pytest

from compute import _capi as capi
from gpu_util import alloc, download, upload
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu
lib = capi.lib


def _both(T, x, rows, cols):
    b = upload(T, x)
    outs = {}
    for name, fn in (("alloc", lib.cs_reduce_rows), ("noalloc", lib.cs_reduce_rows_noalloc)):
        rs, rm, tot, tot64 = alloc(T, rows), alloc(T, rows), alloc(T, 1), alloc(T, 1, np.float64)
        capi.check(fn(T.stream, b.ptr, rows, cols, rs.ptr, rm.ptr, tot.ptr, tot64.ptr))
        outs[name] = (download(T, rs, rows), download(T, rm, rows), download(T, tot, 1),
                      download(T, tot64, 1, np.float64))
        for h in (rs, rm, tot, tot64):
            h.release()
    b.release()
    return outs


def test_noalloc_matches_the_allocating_entry_point(gpu_tensors):
    T = gpu_tensors
    for rows, cols, seed in ((16384, 4096, 3001), (1000, 301, 77)):  # the second call is smaller
        print("noalloc rows %d cols %d" % (rows, cols), flush=True)
        x = inputs.heavy_tails(inputs.normal(rows * cols, seed), seed + 1).reshape(rows, cols)
        outs = _both(T, x, rows, cols)
        for got, want in zip(outs["noalloc"], outs["alloc"]):
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
        s64 = x.astype(np.float64).sum(axis=1)
        sa = np.abs(x.astype(np.float64)).sum(axis=1)
        assert np.all(np.abs(outs["noalloc"][0] - s64) <= 4 * fp64.EPS32 * sa)
        assert np.array_equal(outs["noalloc"][1], x.max(axis=1))
        assert abs(float(outs["noalloc"][3][0]) - s64.sum()) <= 1e-12 * sa.sum()


def test_noalloc_limits(gpu_tensors):
    T = gpu_tensors
    rows = (1 << 20) + 1
    b = alloc(T, rows)
    rs, tot = alloc(T, rows), alloc(T, 1)
    with pytest.raises(capi.InvalidBufferSize):
        capi.check(lib.cs_reduce_rows_noalloc(T.stream, b.ptr, rows, 1, None, None, tot.ptr, None))
    # row sums alone need no partials: any row count
    capi.check(lib.cs_reduce_rows_noalloc(T.stream, b.ptr, rows, 1, rs.ptr, None, None, None))
    T.synchronize()
    # no rows: the total is the empty sum
    capi.check(lib.cs_reduce_rows_noalloc(T.stream, None, 0, 5, None, None, tot.ptr, None))
    assert download(T, tot, 1)[0] == 0.0
    for h in (b, rs, tot):
        h.release()
