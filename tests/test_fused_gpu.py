"""Full-size parity of the fused kernels of the two elementwise workloads:
  a*b+c on [1024,1024] (U[-1,1], seeds 1001-1003), and
  the view chain permute(1,0,2) + bias broadcast + translate(0,0,1), then
  exp(-y*y)*tanh(y), on X [64,256,256] ~ N(0,1) (seed 2001), bias ~ U(-0.5,0.5) (seed 2002).
Tolerances: views and + - * exact (bit-exact against the oracle); each exp/tanh within
2 ulp of its exact value in the precise mode (1 ulp in the accurate mode), so the output lies
within the propagated bound around the FP32 program with correctly rounded operations."""
import numpy as np
import pytest

from oracle import fp64, inputs, refcl

pytestmark = pytest.mark.gpu


def test_c1_fused_multiply_add_bit_exact(gpu_tensors):
    T = gpu_tensors
    a, b, c = inputs.c1_inputs()
    r = T.Tensor(a) * T.Tensor(b) + T.Tensor(c)
    got = r.flatArray()
    want = ((a * b) + c).ravel()  # IEEE FP32, no contraction
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    k = refcl.Kernel(r.closure, r.shape)
    assert np.array_equal(k([a, b, c]).view(np.uint32), got.view(np.uint32))


def _c2(T, X, bias):
    p = T.Tensor(X).permute([1, 0, 2])
    y = p + T.Tensor(bias).broadcast((256, 64, 256)) + p.translate([0, 0, 1])
    return y, T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)


def test_c2_view_chain(gpu_tensors):
    T = gpu_tensors
    X, bias = inputs.c2_inputs()
    y, out = _c2(T, X, bias)
    y_gpu = y.flatArray().reshape(256, 64, 256)
    # views and adds: bit exact against an independent numpy statement
    P = np.transpose(X, (1, 0, 2))
    shifted = np.zeros_like(P)
    shifted[:, :, 1:] = P[:, :, :-1]
    y_np = (P + bias[:, None, None]) + shifted
    assert np.array_equal(y_gpu.view(np.uint32), y_np.view(np.uint32))
    # and against the reference's kernel text run as host C
    ky = refcl.Kernel(y.closure, y.shape)
    assert np.array_equal(ky([X, bias]).reshape(y_np.shape).view(np.uint32), y_gpu.view(np.uint32))
    o_gpu = out.flatArray()
    want = fp64.evaluate(out.closure, out.shape, [X, bias], "fp32").ravel()
    d = fp64.ulp_distance(o_gpu, want)
    print("c2 output ulp: max %d, mean %.4f, > 1ulp: %d, > 2ulp: %d" %
          (d.max(), d.mean(), int((d > 1).sum()), int((d > 2).sum())))
    value, bound = fp64.evaluate_with_bound(out.closure, out.shape, [X, bias], ulps=2.0)
    assert fp64.within_bound(o_gpu, value, bound).all()
    ko = refcl.Kernel(out.closure, out.shape)
    d2 = fp64.ulp_distance(o_gpu, ko([X, bias]))
    print("c2 vs refcl-host (glibc): max %d" % d2.max())
    assert d2.max() <= 4


@pytest.mark.parametrize("mode", ["precise", "accurate"])
def test_math_functions_over_every_float(gpu_tensors, mode):
    """exp, tanh, log and sqrt as generated kernels compute them: every FP32 input (2^32
    of them) against FP64 rounded to FP32, on the device. Precise mode: each within 2 ulp;
    accurate mode: exp and tanh within 1 ulp; sqrt exact in both."""
    import ctypes
    from compute import _capi as capi
    from gpu_util import alloc, download
    T = gpu_tensors
    stats = alloc(T, 3, np.uint64)
    m = {"precise": capi.MATH_PRECISE, "accurate": capi.MATH_ACCURATE}[mode]
    if mode == "precise":
        # the bounded exp (generated where the argument cannot exceed 88) over every x <= 0
        limits = {capi.OP_EXP: 2, capi.OP_TANH: 2, capi.OP_LOG: 2, capi.OP_SQRT: 0, capi.CHECK_EXP_BOUNDED: 2}
    else:
        limits = {capi.OP_EXP: 1, capi.OP_TANH: 1}
    for op, limit in limits.items():
        zero = np.zeros(3, np.uint64)
        capi.check(capi.lib.cs_memcpy_htod(T.stream, ctypes.c_void_p(stats.ptr), zero.ctypes.data_as(ctypes.c_void_p), 24))
        for half in ((1,) if op == capi.CHECK_EXP_BOUNDED else (0, 1)):
            capi.check(capi.lib.cs_math_check(T.stream, op, m, half << 31, 1 << 31, stats.ptr))
        s = download(T, stats, 3, np.uint64)
        x = np.array([s[2]], np.uint32).view(np.float32)[0]
        print("%s op %d: max ulp %d, inputs over 1 ulp %d, worst input %r" % (mode, op, s[0], s[1], x))
        assert s[0] <= limit, (op, s[0], x)
    stats.release()


def test_c2_accurate_mode():
    """The accurate math mode: exp and tanh within 1 ulp, so the output stays within the
    propagated bound with one ulp per transcendental."""
    from compute import Tensors, _capi as capi
    with Tensors(0, math_mode=capi.MATH_ACCURATE) as T:
        X, bias = inputs.c2_inputs()
        y, out = _c2(T, X, bias)
        o = out.flatArray()
        value, bound = fp64.evaluate_with_bound(out.closure, out.shape, [X, bias], ulps=1.0)
        assert fp64.within_bound(o, value, bound).all()
        d = fp64.ulp_distance(o, value)
        print("c2 accurate mode: max %d ulp, > 1ulp: %d" % (d.max(), int((d > 1).sum())))


@pytest.mark.parametrize("opts", [dict(vector_width=1), dict(vector_width=2, items_per_thread=3),
                                  dict(vector_width=8, block_size=64), dict(items_per_thread=8, block_size=1024)])
def test_launch_shapes_agree(opts):
    """Every vector width / items-per-thread / block size computes identical results."""
    from compute import Tensors
    with Tensors(0, **opts) as T:
        X, bias = inputs.c2_inputs()
        y, out = _c2(T, X, bias)
        ref = Tensors(0)
        try:
            _, out_ref = _c2(ref, X, bias)
            assert np.array_equal(out.flatArray().view(np.uint32), out_ref.flatArray().view(np.uint32))
        finally:
            ref.close()


@pytest.mark.parametrize("opts", [dict(vector_width=4), dict(vector_width=8, items_per_thread=2),
                                  dict(vector_width=4, items_per_thread=3, block_size=64)])
def test_shifted_rows_share_chunks(opts):
    """Translates along the innermost dimension read aligned 16-byte chunks of the source row,
    shared between views of one leaf that differ only in the shift; every element keeps the
    padding rule of its own view. Bit-exact with the oracle for shifts inside one chunk,
    across chunks and beyond the row."""
    from compute import Tensors
    with Tensors(0, **opts) as T:
        for shape in [(3, 8, 16), (2, 5, 24), (64,)]:
            n = int(np.prod(shape))
            x = inputs.normal(n, 31).reshape(shape)
            t = T.Tensor(x, padding=7.5)
            for k1, k2 in [(0, 1), (-1, 3), (4, -4), (5, -9), (17, 0), (-2, -3), (-70, 2)]:
                o1 = [0] * (len(shape) - 1) + [k1]
                o2 = [0] * (len(shape) - 1) + [k2]
                e = t.translate(o1) * t + t.translate(o2) - t
                want = fp64.evaluate(e.closure, e.shape, [x], "fp32").ravel()
                got = e.flatArray()
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (opts, shape, k1, k2)


def test_edges_and_odd_shapes(gpu_tensors):
    T = gpu_tensors
    for shape in [(1,), (3,), (5, 7), (2, 3, 5, 7), (1, 1, 1, 1, 1, 3)]:
        n = int(np.prod(shape))
        x = inputs.normal(n, 5).reshape(shape)
        t = T.Tensor(x)
        off = [1] + [0] * (len(shape) - 1)
        e = t.translate(off) * t + t
        want = fp64.evaluate(e.closure, e.shape, [x], "fp32").ravel()
        assert np.array_equal(e.flatArray().view(np.uint32), want.view(np.uint32)), shape
    empty = T.Tensor(np.zeros((0, 4), np.float32))
    assert (empty + empty).flatArray().size == 0


SPECIAL = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-45, -1e-45, 1e-40, 1.17549435e-38, 3.4e38, -3.4e38,
                    1.0, -1.0, 0.5, 88.7, 89.0, -87.0, -104.0, -110.0, 9.5, 20.0, 1e-7, 2.0, 7.0, -3.0],
                   np.float32)


def _same(got, want):
    """Bit-exact, with every NaN equal to every NaN (payloads are not specified)."""
    g, w = got.ravel(), want.ravel()
    nan = np.isnan(w)
    return np.array_equal(np.isnan(g), nan) and np.array_equal(g[~nan].view(np.uint32), w[~nan].view(np.uint32))


@pytest.mark.parametrize("op", ["+", "-", "*", "/", "%", "min", "max"])
def test_special_values_binary(gpu_tensors, op):
    """Every pair of special values (NaN, infinities, signed zeros, subnormals, the extremes)
    through a fused kernel with a broadcast view on each side: IEEE FP32 results bit-exact
    (fmodf for %, fminf/fmaxf for min/max, as OpenCL C defines them), so no contraction,
    flush or fast-math rewrite is present."""
    T = gpu_tensors
    n = SPECIAL.size
    a = T.Tensor(SPECIAL.reshape(n, 1)).broadcast((n, n))
    b = T.Tensor(SPECIAL.reshape(1, n)).broadcast((n, n))
    f = {"+": lambda x, y: x + y, "-": lambda x, y: x - y, "*": lambda x, y: x * y, "/": lambda x, y: x / y,
         "%": lambda x, y: x % y, "min": T.Tensor.min, "max": T.Tensor.max}[op]
    e = f(a, b)
    want = fp64.evaluate(e.closure, e.shape, [SPECIAL.reshape(n, 1), SPECIAL.reshape(1, n)], "fp32")
    got = e.flatArray().reshape(n, n)
    if op in ("min", "max"):
        # fmin/fmax of +0 and -0 may return either zero
        zeros = (want == 0) & (got == 0)
        got, want = np.where(zeros, 0.0, got).astype(np.float32), np.where(zeros, 0.0, want).astype(np.float32)
    assert _same(got, want), op


@pytest.mark.parametrize("op", ["neg", "abs", "sqrt", "exp", "tanh", "log"])
def test_special_values_unary(gpu_tensors, op):
    """Unary ops on the special values: neg, abs and sqrt bit-exact; exp, tanh and log exact
    where the result is special (NaN, infinities, zero, +-1) and within 2 ulp elsewhere."""
    T = gpu_tensors
    x = T.Tensor(SPECIAL)
    e = {"neg": lambda t: -t, "abs": T.Tensor.abs, "sqrt": T.Tensor.sqrt, "exp": T.Tensor.exp,
         "tanh": T.Tensor.tanh, "log": T.Tensor.log}[op](x)
    want = fp64.evaluate(e.closure, e.shape, [SPECIAL], "fp32")
    got = e.flatArray()
    if op in ("neg", "abs", "sqrt"):
        assert _same(got, want), op
        return
    special = np.isnan(want) | np.isinf(want) | (want == 0) | (np.abs(want) == 1)
    assert _same(got[special], want[special]), (op, SPECIAL[special], got[special], want[special])
    d = fp64.ulp_distance(got[~special], want[~special])
    assert d.max() <= 2, (op, SPECIAL[~special][d > 2])


def test_output_near_the_index_limit():
    """An output of 2^31 - 2^20 elements (8 GiB), just under the 32-bit index limit that the
    generated kernels share with the reference's (an int global id): y = x*2 + 1 over a
    U(-1,1) leaf generated on the device, three 1 Mi-element windows (first, middle, last)
    bit-exact against the oracle's sequence."""
    import ctypes
    from compute import Tensors, _capi as capi
    from compute import tensors as tz
    with Tensors(0) as T:
        n = (1 << 31) - (1 << 20)
        buf = T._alloc(n * 4)
        capi.check(capi.lib.cs_fill_uniform(T.stream, buf.ptr, n, 77, -1.0, 1.0))
        x = tz.CachedTensor(T, (n,), 0.0, buf)
        y = (x * T.Tensor.fill(2.0, (n,)) + T.Tensor.fill(1.0, (n,))).doCache()
        m = 1 << 20
        for start in (0, n // 2, n - m):
            got = np.empty(m, np.float32)
            capi.check(capi.lib.cs_buffer_read(T.stream, y.buffer.handle, start * 4,
                                               got.ctypes.data_as(ctypes.c_void_p), m * 4))
            xs = inputs.uniform(m, 77, start=start)
            want = xs * np.float32(2.0) + np.float32(1.0)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), start
        y.release()
        x.release()


def test_large_body_keeps_its_workgroup_size(gpu_tensors):
    """An expression of more than 256 nodes is compiled for its own workgroup size (so that a
    thread may hold up to 512 registers): launching it with a larger block is an error, with
    its own size or smaller it computes the same result."""
    import ctypes
    from functools import reduce
    from compute import _capi as capi
    T = gpu_tensors
    x = inputs.normal(64 * 64, 71).reshape(64, 64)
    t = T.Tensor(x)
    terms = [t.translate([0, k % 7 - 3]) * T.Tensor.fill(float(k + 1), (64, 64)) for k in range(150)]
    e = reduce(lambda a, b: a + b, terms)
    want = e.flatArray()
    f, leaves, _ = T.compile(e.closure, e.shape)
    try:
        out = T._alloc(64 * 64 * 4)
        src = leaves[0].id._buffer({})
        ptrs = (ctypes.c_void_p * 1)(src.ptr)
        with pytest.raises(capi.InvalidWorkGroupSize):
            capi.check(capi.lib.cs_fused_launch_block(T.stream, f, ptrs, 1, out.ptr, 512))
        for block in (256, 128):
            capi.check(capi.lib.cs_fused_launch_block(T.stream, f, ptrs, 1, out.ptr, block))
            got = np.empty(64 * 64, np.float32)
            capi.check(capi.lib.cs_memcpy_dtoh(T.stream, got.ctypes.data_as(ctypes.c_void_p),
                                               ctypes.c_void_p(out.ptr), got.nbytes))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), block
        out.release()
        src.release()
    finally:
        capi.lib.cs_fused_release(f)
