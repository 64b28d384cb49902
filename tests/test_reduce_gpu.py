"""Reductions against the FP64 oracle: sum within 4*eps32*sum|x|, max exact.
Full size: A [16384,4096] ~ N(0,1) with 1% of entries x1e4 (seeds 3001/3002)."""
import ctypes

import numpy as np
import pytest

import oracle
from compute import _capi as capi
from gpu_util import alloc, download, upload
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu


def test_sum_known_answer(gpu_tensors):
    assert gpu_tensors.Tensor.fill(15625.0, (8, 8)).sum().toString() == "1000000.0"


@pytest.mark.parametrize("n", [0, 1, 3, 255, 256, 1000, 4097, 1 << 20, (1 << 20) + 13])
def test_full_sum_sizes(gpu_tensors, n):
    T = gpu_tensors
    x = inputs.heavy_tails(inputs.normal(n, 77), 78) if n else np.zeros(0, np.float32)
    b = upload(T, x)
    out = alloc(T, 1)
    out64 = alloc(T, 1, np.float64)
    capi.check(capi.lib.cs_reduce_sum(T.stream, b.ptr, n, out.ptr, out64.ptr))
    s = download(T, out, 1)[0]
    s64 = download(T, out64, 1, np.float64)[0]
    exact = x.astype(np.float64).sum()
    assert abs(float(s) - exact) <= fp64.sum_bound(x) + abs(exact) * 2 ** -24
    assert abs(s64 - exact) <= 1e-12 * max(np.abs(x).sum(), 1.0)
    if n:
        capi.check(capi.lib.cs_reduce_max(T.stream, b.ptr, n, out.ptr))
        assert download(T, out, 1)[0] == x.max()
    for h in (b, out, out64):
        h.release()


def test_max_ignores_nan(gpu_tensors):
    T = gpu_tensors
    x = np.array([np.nan, -3.0, 2.0, np.nan, -7.0], np.float32)
    t = T.Tensor(x)
    assert t.maxValue().flatArray()[0] == 2.0


def test_rows_full_size(gpu_tensors):
    T = gpu_tensors
    rows, cols = 16384, 4096
    a = inputs.c3_input(rows, cols)
    b = upload(T, a)
    rs, rm, tot = alloc(T, rows), alloc(T, rows), alloc(T, 1)
    tot64 = alloc(T, 1, np.float64)
    for _ in range(2):  # the ticket resets itself between calls
        capi.check(capi.lib.cs_reduce_rows(T.stream, b.ptr, rows, cols, rs.ptr, rm.ptr, tot.ptr, tot64.ptr))
    s64, m64, sa, total = oracle.row_stats64(a)
    got_s = download(T, rs, rows)
    got_m = download(T, rm, rows)
    err = np.abs(got_s.astype(np.float64) - s64)
    bound = 4 * fp64.EPS32 * sa
    print("row sum error / bound: max %.3g" % (err / bound).max())
    assert np.all(err <= bound)
    assert np.array_equal(got_m, m64)
    t = float(download(T, tot, 1)[0])
    t64 = float(download(T, tot64, 1, np.float64)[0])
    assert abs(t - total) <= 4 * fp64.EPS32 * sa.sum()
    assert abs(t64 - total) <= 1e-12 * sa.sum()
    # the left fold the reference would build (split(1).reduce(_ + _)) sits inside the same bound
    fold_s, _ = oracle.row_fold(a)
    assert np.all(np.abs(fold_s.astype(np.float64) - s64) <= cols * fp64.EPS32 * sa)
    for h in (b, rs, rm, tot, tot64):
        h.release()


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (7, 1023), (129, 4099), (1000, 3), (3, 300001),
                                   (1, 8), (300, 1024), (1031, 68), (2049, 4)])
def test_rows_and_cols_ragged(gpu_tensors, shape):
    T = gpu_tensors
    rows, cols = shape
    a = inputs.normal(rows * cols, 9).reshape(shape)
    t = T.Tensor(a)
    s64, m64, sa, _ = oracle.row_stats64(a)
    assert np.all(np.abs(t.rowSum().flatArray() - s64) <= 4 * fp64.EPS32 * sa + 1e-30)
    assert np.array_equal(t.rowMax().flatArray(), m64)
    cs = t.colSum().flatArray()
    c64 = a.astype(np.float64).sum(axis=0)
    assert np.all(np.abs(cs - c64) <= 4 * fp64.EPS32 * np.abs(a).astype(np.float64).sum(axis=0) + 1e-30)


def test_total_of_many_rows(gpu_tensors):
    """The FP64 combine over one partial per row: row counts that fill the eight loads in
    flight per thread exactly, leave an odd partial, and take several rounds of loads (131071
    and 300001 rows). Each total is computed twice, each time after a small sum on the same
    stream, and must give the same bits both times."""
    T = gpu_tensors
    cols = 4
    small = upload(T, np.ones(10, np.float32))
    small_out = alloc(T, 1)
    for rows in (2048, 2049, 131071, 300001):
        a = inputs.heavy_tails(inputs.normal(rows * cols, 401), 402).reshape(rows, cols)
        b = upload(T, a)
        tot, tot64 = alloc(T, 2), alloc(T, 2, np.float64)
        for k in range(2):
            capi.check(capi.lib.cs_reduce_sum(T.stream, small.ptr, 10, small_out.ptr, None))
            capi.check(capi.lib.cs_reduce_rows(T.stream, b.ptr, rows, cols, None, None, tot.ptr + 4 * k,
                                               tot64.ptr + 8 * k))
        _, _, sa, total = oracle.row_stats64(a)
        t = download(T, tot, 2)
        t64 = download(T, tot64, 2, np.float64)
        print("rows %d: total %.17g, oracle %.17g" % (rows, t64[0], total), flush=True)
        assert t64[0] == t64[1] and t[0] == t[1]
        assert abs(float(t64[0]) - total) <= 1e-12 * sa.sum()
        assert abs(float(t[0]) - total) <= 4 * fp64.EPS32 * sa.sum()
        assert download(T, small_out, 1)[0] == 10.0
        for h in (b, tot, tot64):
            h.release()
    for h in (small, small_out):
        h.release()


@pytest.mark.parametrize("kind", ["sum", "max", "rows"])
def test_combine_reads_every_new_partial(gpu_tensors, kind):
    """The combine adds the partials that this call's workgroups wrote. Two inputs of the same
    shape alternate over 200 back-to-back calls with no host wait between them, so each call's
    partials overwrite the previous call's in the same scratch: a combine that read one stale
    partial would mix the two inputs. Then the
    same pair of launches captured in a graph, replayed five times with other reductions on
    the stream between replays. Every result must carry its own input's bits."""
    T = gpu_tensors
    rows, cols = 4096, 1024
    xs = [inputs.heavy_tails(inputs.normal(rows * cols, 501 + k), 503 + k).reshape(rows, cols) for k in range(2)]
    bufs = [upload(T, x) for x in xs]
    launches = 200
    outs = alloc(T, launches)
    outs64 = alloc(T, launches, np.float64)
    f4, f8 = ctypes.sizeof(ctypes.c_float), ctypes.sizeof(ctypes.c_double)

    def launch(i, k, slot):
        o, o64 = outs.ptr + slot * f4, outs64.ptr + slot * f8
        if kind == "sum":
            capi.check(capi.lib.cs_reduce_sum(T.stream, bufs[k].ptr, rows * cols, o, o64))
        elif kind == "max":
            capi.check(capi.lib.cs_reduce_max(T.stream, bufs[k].ptr, rows * cols, o))
        else:
            capi.check(capi.lib.cs_reduce_rows(T.stream, bufs[k].ptr, rows, cols, None, None, o, o64))

    for i in range(launches):
        launch(i, i % 2, i)
    got = download(T, outs, launches)
    got64 = download(T, outs64, launches, np.float64)
    for k in range(2):
        mine = got[k::2]
        assert np.array_equal(mine.view(np.uint32), np.full(len(mine), mine[0], np.float32).view(np.uint32)), k
        if kind == "max":
            assert mine[0] == xs[k].max()
        else:
            assert np.all(got64[k::2] == got64[k])
            exact = xs[k].astype(np.float64).sum()
            assert abs(got64[k] - exact) <= 1e-12 * np.abs(xs[k]).astype(np.float64).sum()
    assert got[0] != got[1]
    first = got[:2].copy()
    # the same two launches in a graph, replayed with eager reductions between the replays
    capi.check(capi.lib.cs_graph_begin_capture(T.stream))
    launch(0, 0, 0)
    launch(1, 1, 1)
    g = ctypes.c_void_p()
    capi.check(capi.lib.cs_graph_end_capture(T.stream, ctypes.byref(g)))
    for r in range(5):
        capi.check(capi.lib.cs_graph_launch(T.stream, g))
        launch(2 + 2 * r, 1, 2 + 2 * r)
        launch(3 + 2 * r, 0, 3 + 2 * r)
        replay = download(T, outs, 2)
        assert np.array_equal(replay.view(np.uint32), first.view(np.uint32)), r
    capi.lib.cs_graph_release(g)
    eager = download(T, outs, 12)
    assert np.array_equal(eager[2:12:2].view(np.uint32), np.full(5, first[1], np.float32).view(np.uint32))
    assert np.array_equal(eager[3:12:2].view(np.uint32), np.full(5, first[0], np.float32).view(np.uint32))
    for h in bufs + [outs, outs64]:
        h.release()


@pytest.mark.parametrize("kind", ["sum", "max", "rows"])
def test_combine_is_deterministic(gpu_tensors, kind):
    """Block partials are combined in FP64 in a fixed order: 300 back-to-back launches (each
    followed by a one-block reduction on the same stream) must all give the first launch's
    bits and the FP64 oracle's value."""
    T = gpu_tensors
    rows, cols = 4096, 1024
    a = inputs.heavy_tails(inputs.normal(rows * cols, 301), 302).reshape(rows, cols)
    b = upload(T, a)
    launches = 300
    outs = alloc(T, launches)
    outs64 = alloc(T, launches, np.float64)
    small = upload(T, a[0, :100].copy())  # one block
    small_out = alloc(T, launches)
    f4, f8 = ctypes.sizeof(ctypes.c_float), ctypes.sizeof(ctypes.c_double)
    for i in range(launches):
        o, o64 = outs.ptr + i * f4, outs64.ptr + i * f8
        if kind == "sum":
            capi.check(capi.lib.cs_reduce_sum(T.stream, b.ptr, rows * cols, o, o64))
        elif kind == "max":
            capi.check(capi.lib.cs_reduce_max(T.stream, b.ptr, rows * cols, o))
        else:
            capi.check(capi.lib.cs_reduce_rows(T.stream, b.ptr, rows, cols, None, None, o, o64))
        capi.check(capi.lib.cs_reduce_sum(T.stream, small.ptr, 100, small_out.ptr + i * f4, None))
    got = download(T, outs, launches)
    assert np.array_equal(got.view(np.uint32), np.full(launches, got[0], np.float32).view(np.uint32))
    small_got = download(T, small_out, launches)
    assert np.all(small_got == small_got[0])
    assert abs(float(small_got[0]) - a[0, :100].astype(np.float64).sum()) <= fp64.sum_bound(a[0, :100]) * 2
    if kind == "max":
        assert got[0] == a.max()
    else:
        got64 = download(T, outs64, launches, np.float64)
        assert np.all(got64 == got64[0])
        exact = a.astype(np.float64).sum()
        assert abs(got64[0] - exact) <= 1e-12 * np.abs(a).astype(np.float64).sum()
    for h in (b, outs, outs64, small, small_out):
        h.release()


@pytest.mark.parametrize("shape,d,op", [((64, 300), 1, "+"), ((64, 300), 1, "max"), ((5, 7, 33), 2, "+"),
                                        ((40, 50), 0, "+"), ((3, 20, 17), 1, "+"), ((40, 50), 0, "max")])
def test_split_fold_is_an_axis_reduction(shape, d, op):
    """t.split(d).reduce(+ or max), the reference's way to write axis reductions: recognised
    and run as one reduction (sums within 4*eps32*sum|x| of FP64, maxima exact); the max over
    a non-innermost dimension stays a fused kernel, bit-exact with the fold."""
    import functools
    from compute import Tensors
    with Tensors(0) as T:
        n = int(np.prod(shape))
        x = inputs.heavy_tails(inputs.normal(n, 91), 92).reshape(shape)
        t = T.Tensor(x).doCache()
        slices = list(t.split(d))
        fold = functools.reduce((lambda a, b: a + b) if op == "+" else T.Tensor.max, slices)
        before = T.launches
        got = fold.flatArray().reshape(tuple(s for i, s in enumerate(shape) if i != d))
        x64 = x.astype(np.float64)
        if op == "+":
            assert T.launches - before == 1
            exact = x64.sum(axis=d)
            bound = 4 * np.finfo(np.float32).eps * np.abs(x64).sum(axis=d)
            assert (np.abs(got - exact) <= bound).all()
        else:
            want = x.max(axis=d)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
            if d != len(shape) - 1:
                ref = fp64.evaluate(fold.closure, fold.shape, [x], "fp32").reshape(got.shape)
                assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        t.release()


def test_rows_of_empty_matrices(gpu_tensors):
    """No rows: the total is 0 (written, not left as it was); no columns: row sums 0."""
    T = gpu_tensors
    tot, tot64 = upload(T, np.array([5.0], np.float32)), alloc(T, 1, np.float64)
    capi.check(capi.lib.cs_memcpy_htod(T.stream, ctypes.c_void_p(tot64.ptr),
                                       np.array([5.0]).ctypes.data_as(ctypes.c_void_p), 8))
    capi.check(capi.lib.cs_reduce_rows(T.stream, None, 0, 7, None, None, tot.ptr, tot64.ptr))
    assert download(T, tot, 1)[0] == 0.0 and download(T, tot64, 1, np.float64)[0] == 0.0
    rs = upload(T, np.full(3, 5.0, np.float32))
    capi.check(capi.lib.cs_reduce_rows(T.stream, None, 3, 0, rs.ptr, None, tot.ptr, None))
    assert np.array_equal(download(T, rs, 3), np.zeros(3, np.float32)) and download(T, tot, 1)[0] == 0.0
    for h in (tot, tot64, rs):
        h.release()


def test_fused_sum_of_an_inline_tensor():
    """`(a*b + c).sum` runs as one kernel that computes the elements and adds them (no output
    written): within 4*eps32*sum|x| of the FP64 sum of the FP32 element values, within
    8*eps32*sum|x| of the unfused sum, the same bits on every run, one launch."""
    from compute import Tensors
    a, b, c = inputs.c1_inputs()
    with Tensors(0) as T, Tensors(0, fused_sums=False) as U:
        ta, tb, tc = (T.Tensor(x).doCache() for x in (a, b, c))
        e = ta * tb + tc
        before = T.launches
        s1 = e.sum().flatArray()[0]
        assert T.launches - before == 1
        s2 = e.sum().flatArray()[0]
        assert np.float32(s1).view(np.uint32) == np.float32(s2).view(np.uint32)
        vals = ((a * b) + c).astype(np.float64)
        bound = 4 * fp64.EPS32 * np.abs(vals).sum()
        assert abs(float(s1) - vals.sum()) <= bound
        ua, ub, uc = (U.Tensor(x) for x in (a, b, c))
        s_unfused = (ua * ub + uc).sum().flatArray()[0]
        assert abs(float(s1) - float(s_unfused)) <= 2 * bound
        for t in (ta, tb, tc):
            t.release()


def test_fused_sum_of_views_and_edges(gpu_tensors):
    """The view chain with exp and tanh, a broadcast fill, an empty tensor: fused sums agree
    with the sum of the element-wise kernel's own output."""
    T = gpu_tensors
    X, bias = inputs.c2_inputs()
    p = T.Tensor(X).permute([1, 0, 2])
    y = p + T.Tensor(bias).broadcast((256, 64, 256)) + p.translate([0, 0, 1])
    out = T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)
    vals = out.flatArray().astype(np.float64)
    s = out.sum().flatArray()[0]
    assert abs(float(s) - vals.sum()) <= 4 * fp64.EPS32 * np.abs(vals).sum()
    assert T.Tensor.fill(15625.0, (8, 8)).sum().toString() == "1000000.0"
    empty = T.Tensor(np.zeros((0, 5), np.float32))
    assert (empty + empty).sum().flatArray()[0] == 0.0


def test_sum_kernel_is_not_an_elementwise_kernel(gpu_tensors):
    T = gpu_tensors
    x = T.Tensor(np.ones((4, 4), np.float32))
    f, leaves = T.compile_sum((x + x).closure, (4, 4))
    try:
        out, leaf = alloc(T, 16), leaves[0].id._buffer({})
        ptrs = (ctypes.c_void_p * 1)(leaf.ptr)
        with pytest.raises(capi.InvalidKernelArgs):
            capi.check(capi.lib.cs_fused_launch(T.stream, f, ptrs, 1, out.ptr))
        out.release()
        leaf.release()
    finally:
        capi.lib.cs_fused_release(f)


def _same_class(got, want):
    """Infinities with the same sign, NaN where want is NaN; finite values compared by the caller."""
    got, want = np.asarray(got, np.float32), np.asarray(want, np.float32)
    return (np.array_equal(np.isnan(got), np.isnan(want)) and
            np.array_equal(np.isposinf(got), np.isposinf(want)) and
            np.array_equal(np.isneginf(got), np.isneginf(want)))


SPECIALS = [[1.0, np.inf, 2.0], [-np.inf, 1.0], [np.inf, -np.inf], [3e38, 3e38], [3e38, 3e38, -1.0],
            [np.nan, 1.0], [-3e38, -3e38]]


@pytest.mark.parametrize("spread", [False, True])
@pytest.mark.parametrize("case", range(len(SPECIALS)))
def test_sums_of_non_finite_values(gpu_tensors, case, spread):
    """A compensated sum must not turn an infinity or an FP32 overflow into NaN: the full sum,
    rowSum, colSum, split(d).reduce(_ + _) and the fused sum give what numpy's FP32 sum gives
    ([1, inf, 2] -> inf, [3e38, 3e38] -> inf, [inf, -inf] -> NaN). With `spread`, the values sit
    apart among 300,000 zeros, so they reach different threads and workgroups."""
    import functools
    T = gpu_tensors
    vals = np.array(SPECIALS[case], np.float32)
    n = 300000 if spread else len(vals)
    x = np.zeros(n, np.float32)
    where = np.linspace(0, n - 1, len(vals)).astype(np.int64) if spread else np.arange(len(vals))
    x[where] = vals
    with np.errstate(over="ignore", invalid="ignore"):
        want = np.float32(x.sum(dtype=np.float32))
    t = T.Tensor(x).doCache()
    zero = T.Tensor(np.zeros(n, np.float32)).doCache()
    for name, got in (("sum", t.sum().flatArray()[0]), ("fused", (t + zero).sum().flatArray()[0])):
        assert _same_class(got, want), (name, got, want)
        if np.isfinite(want):
            assert abs(float(got) - x.astype(np.float64).sum()) <= 4 * fp64.EPS32 * np.abs(x.astype(np.float64)).sum()
    # rows: the case in row 1 of three, columns: the case in column 1 of three
    m = np.zeros((3, n), np.float32)
    m[1] = x
    tm = T.Tensor(m).doCache()
    with np.errstate(over="ignore", invalid="ignore"):
        want_rows = m.sum(axis=1, dtype=np.float32)
        want_cols = m.T.copy().sum(axis=0, dtype=np.float32)
    rs = tm.rowSum().flatArray()
    assert _same_class(rs, want_rows), (rs, want_rows)
    if not spread:  # a fold over 300,000 slices is not a program anyone writes
        fold = functools.reduce(lambda a, b: a + b, list(tm.split(1)))
        fs = fold.flatArray()
        assert _same_class(fs, want_rows), (fs, want_rows)
    tc = T.Tensor(m.T.copy()).doCache()
    cs = tc.colSum().flatArray()
    assert _same_class(cs, want_cols), (cs, want_cols)
    for h in (t, zero, tm, tc):
        h.release()
