"""The fast math mode, the counterpart of -cl-unsafe-math-optimizations that the reference's
`gpu` and `cpu` objects compile with (gpu.scala:15-17, cpu.scala:105, OpenCL.scala:1131-1135):
contraction, reassociation, reciprocals and no signed zeros, but infinities and NaN keep their
meaning. Checked against the FP64 oracle's fast bound (oracle/fp64.py: one ulp of the working
magnitude per + - *, two per /, exp within FAST_EXP_ULPS[0] + FAST_EXP_ULPS[1]*|x| ulp, log
within FAST_LOG_ABS absolute plus FAST_LOG_ULPS ulp) and against the reference's kernel text run
as host C in its fast mode (refcl-host "fast")."""
import ctypes

import numpy as np
import pytest

from compute import _capi as capi
from gpu_util import alloc, download
from oracle import fp64, inputs, refcl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fast():
    from compute import Tensors
    T = Tensors(0, math_mode=capi.MATH_FAST)
    yield T
    T.close()


def _binade_allowance_exp(e):
    """Largest ulp error the fast exp model allows in the binade 2^e <= |x| < 2^(e+1)."""
    return fp64.FAST_EXP_ULPS[0] + fp64.FAST_EXP_ULPS[1] * 2.0 ** (e + 1) + 0.5


@pytest.mark.parametrize("op", ["exp", "log"])
def test_fast_functions_over_every_float(fast, op):
    """__expf and __logf, the functions the fast mode emits, over all 2^32 inputs, one binade
    (2^23 inputs of one sign and exponent) per check, against FP64 rounded to FP32: exp within
    the model's allowance at the binade's largest |x| wherever its result is a normal number
    (below 2^-126 it may flush to zero, which the sampled test checks against FAST_FLUSH);
    log within FAST_LOG_ULPS ulp everywhere, NaN for negative arguments."""
    T = fast
    stats = alloc(T, 3, np.uint64)
    zero = np.zeros(3, np.uint64)
    code = {"exp": capi.OP_EXP, "log": capi.OP_LOG}[op]
    flush_at = int(np.array([np.log(2.0 ** -126)], np.float32).view(np.uint32)[0])  # exp's result there is 2^-126
    worst = {}
    for sign in (0, 1):
        for e in range(256):
            start = (sign << 31) | (e << 23)
            count = 1 << 23
            if op == "exp" and sign == 1 and start < flush_at < start + count:
                count = flush_at - start  # arguments whose exp is a normal number
            elif op == "exp" and sign == 1 and start >= flush_at and e < 255:
                continue  # every result below 2^-126
            capi.check(capi.lib.cs_memcpy_htod(T.stream, ctypes.c_void_p(stats.ptr),
                                               zero.ctypes.data_as(ctypes.c_void_p), 24))
            capi.check(capi.lib.cs_math_check(T.stream, code, capi.MATH_FAST, start, count, stats.ptr))
            s = download(T, stats, 3, np.uint64)
            worst[(sign, e - 127)] = int(s[0])
            x = np.array([s[2]], np.uint32).view(np.float32)[0]
            if op == "exp":
                if e == 255:
                    assert s[0] == 0, (sign, x)  # inf and NaN exact
                    continue
                assert s[0] <= _binade_allowance_exp(e - 127), (sign, e - 127, int(s[0]), x)
            else:
                if sign == 1 and e > 0:
                    assert s[0] == 0, x  # negative arguments: NaN (the zero binade holds -0)
                else:
                    assert s[0] <= fp64.FAST_LOG_ULPS, (sign, e - 127, int(s[0]), x)
    print("fast %s: worst ulp per binade (sign, exponent): %s" % (op, sorted(worst.items(), key=lambda kv: -kv[1])[:8]))
    stats.release()


def _sample_floats(per_binade=256):
    """Finite FP32 values of both signs: `per_binade` evenly spaced mantissas per exponent."""
    m = (np.arange(per_binade, dtype=np.uint64) * ((1 << 23) // per_binade) + 12345) & ((1 << 23) - 1)
    e = np.arange(255, dtype=np.uint64)
    bits = (e[:, None] << np.uint64(23)) | m[None, :]
    bits = np.concatenate([bits.ravel(), bits.ravel() | np.uint64(1 << 31)]).astype(np.uint32)
    return bits.view(np.float32)


@pytest.mark.parametrize("op", ["exp", "log", "sqrt", "tanh"])
def test_fast_generated_kernels_sampled(fast, op):
    """The generated fast-mode kernels themselves (compiled with the mode's flags), on 256
    mantissas of every finite binade of both signs plus a million points in [0.5, 2): within
    the fast bound around the correctly rounded FP32 value; NaN and infinities where it has
    them."""
    T = fast
    x = np.concatenate([_sample_floats(), inputs.uniform(1 << 20, 41, 0.5, 2.0)])
    fn = {"exp": T.Tensor.exp, "log": T.Tensor.log, "sqrt": T.Tensor.sqrt, "tanh": T.Tensor.tanh}[op]
    e = fn(T.Tensor(x))
    got = e.flatArray()
    value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [x], mode="fast")
    if op == "sqrt":
        # a subnormal argument may be flushed to a zero of its sign first: sqrt(-tiny) is then
        # -0 instead of NaN
        flushed = (x < 0) & (x > -fp64.FAST_FLUSH)
        assert np.all(np.isnan(got[flushed]) | (got[flushed] == 0))
        got, value, bound, x = got[~flushed], value[~flushed], bound[~flushed], x[~flushed]
    ok = fp64.within_bound(got, value, bound)
    assert ok.all(), (op, x[~ok][:5], got[~ok][:5], value[~ok][:5])
    assert np.array_equal(np.isnan(got), np.isnan(value)) and np.array_equal(np.isinf(got), np.isinf(value))
    d = fp64.ulp_distance(got, value)
    print("fast %s: max %d ulp from the correctly rounded value" % (op, d.max()))


SPECIAL = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 7.0, -3.0, 1e30, -1e-30], np.float32)


def _unsigned_zero(x):
    x = np.asarray(x, np.float32)
    return np.where(x == 0, np.float32(0.0), x).astype(np.float32)


def _same_class(got, want):
    return (np.array_equal(np.isnan(got), np.isnan(want)) and np.array_equal(np.isposinf(got), np.isposinf(want))
            and np.array_equal(np.isneginf(got), np.isneginf(want)))


@pytest.mark.parametrize("op", ["+", "-", "*", "/", "%", "min", "max"])
def test_fast_special_values_binary(fast, op):
    """Every pair of special values through a fast-mode kernel and through refcl-host's fast
    mode: NaN and infinities (with their signs) exactly where IEEE FP32 has them, finite
    results within the fast bound (zeros compared without their sign: no signed zeros)."""
    T = fast
    n = SPECIAL.size
    A, B = SPECIAL.reshape(n, 1), SPECIAL.reshape(1, n)
    a = T.Tensor(A).broadcast((n, n))
    b = T.Tensor(B).broadcast((n, n))
    f = {"+": lambda x, y: x + y, "-": lambda x, y: x - y, "*": lambda x, y: x * y, "/": lambda x, y: x / y,
         "%": lambda x, y: x % y, "min": T.Tensor.min, "max": T.Tensor.max}[op]
    e = f(a, b)
    value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [A, B], mode="fast")
    got = e.flatArray().reshape(n, n)
    host = refcl.Kernel(e.closure, e.shape, "fast")([A, B]).reshape(n, n) if op != "%" else None
    for name, r in (("gpu", got), ("refcl fast", host)):
        if r is None:
            continue
        assert _same_class(r, value), (op, name)
        fin = np.isfinite(value)
        assert fp64.within_bound(_unsigned_zero(r[fin]), _unsigned_zero(value[fin]), bound[fin]).all(), (op, name)


def test_fast_special_values_unary(fast):
    T = fast
    for name, fn in (("exp", T.Tensor.exp), ("log", T.Tensor.log), ("sqrt", T.Tensor.sqrt),
                     ("tanh", T.Tensor.tanh), ("neg", lambda t: -t), ("abs", T.Tensor.abs)):
        e = fn(T.Tensor(SPECIAL))
        value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [SPECIAL], mode="fast")
        got = e.flatArray()
        assert _same_class(got, value), name
        fin = np.isfinite(value)
        assert fp64.within_bound(_unsigned_zero(got[fin]), _unsigned_zero(value[fin]), bound[fin]).all(), name


def test_c1_fast_full_size(fast):
    """a*b+c on [1024,1024] in the fast mode: the product may be contracted into the addition,
    so the result is within the fast bound (not bit-exact with the uncontracted program), and
    so is refcl-host's fast mode; the sum of it stays within 4*eps32*sum|x|."""
    T = fast
    a, b, c = inputs.c1_inputs()
    r = T.Tensor(a) * T.Tensor(b) + T.Tensor(c)
    got = r.flatArray()
    value, bound = fp64.evaluate_with_bound(r.closure, r.shape, [a, b, c], mode="fast")
    assert fp64.within_bound(got, value, bound).all()
    fma = (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32).ravel()
    host = refcl.Kernel(r.closure, r.shape, "fast")([a, b, c])
    assert fp64.within_bound(host, value, bound).all()
    plain = ((a * b) + c).ravel()
    print("c1 fast: %d of %d elements differ from the uncontracted program, %d from FMA rounding in FP64; "
          "refcl fast equals the uncontracted program: %s" %
          (int((got != plain).sum()), got.size, int((got != fma).sum()), bool(np.array_equal(host, plain))))
    s = float(r.sum().flatArray()[0])
    vals = got.astype(np.float64)
    assert abs(s - vals.sum()) <= 4 * fp64.EPS32 * np.abs(vals).sum()


def test_c2_fast_full_size(fast):
    """The view chain and exp(-y*y)*tanh(y) on X [64,256,256] in the fast mode: the view sum y
    is bit-exact (adds only, no contraction possible) unless reassociated, and then within the
    fast bound; the output is within the fast bound, as is refcl-host's fast mode."""
    T = fast
    X, bias = inputs.c2_inputs()
    p = T.Tensor(X).permute([1, 0, 2])
    y = p + T.Tensor(bias).broadcast((256, 64, 256)) + p.translate([0, 0, 1])
    out = T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)
    yv, yb = fp64.evaluate_with_bound(y.closure, y.shape, [X, bias], mode="fast")
    y_gpu = y.flatArray()
    assert fp64.within_bound(y_gpu, yv, yb).all()
    o = out.flatArray()
    value, bound = fp64.evaluate_with_bound(out.closure, out.shape, [X, bias], mode="fast")
    assert fp64.within_bound(o, value, bound).all()
    host = refcl.Kernel(out.closure, out.shape, "fast")([X, bias])
    assert fp64.within_bound(host, value, bound).all()
    d = fp64.ulp_distance(o, value)
    dh = fp64.ulp_distance(o, host)
    print("c2 fast: y bit-exact %s; output max %d ulp from the correctly rounded program (mean %.3f), "
          "max %d ulp from refcl-host fast" % (bool(np.array_equal(y_gpu, yv.ravel())), d.max(), d.mean(), dh.max()))


@pytest.mark.parametrize("D,V", [(8, 1), (3, 4)])
def test_convolution_idiom_fast(D, V):
    """The convolution idiom in the fast mode, with its weights read from the LDS table and (at
    one element per vector) a pixel's channels read as one row: every output within the fast
    bound of the FP64 oracle (contracted and reassociated sums of 72 products)."""
    from compute import Tensors, trees
    import spec_cases
    B, H, W, F = 2, 9, 12, 4
    x = inputs.normal(B * H * W * D, 71).reshape(B, H, W, D)
    w = inputs.normal(9 * D * F, 72).reshape(3, 3, D, F)
    b = inputs.normal(F, 73)
    with Tensors(0, math_mode=capi.MATH_FAST, vector_width=V) as T:
        tx, tw, tb = T.Tensor(x), T.Tensor(w), T.Tensor(b)
        out = spec_cases.convolute(T, tx, tw, tb)
        clo = trees.Concatenate([p.closure for p in out._parts])
        got = np.asarray(out.flatArray(), np.float32).reshape(B, H, W, F)
        assert "csmi_k[" in T.kernel_source(clo, (B, H, W))
        data = {id(tx): x, id(tw): w, id(tb): b}
    for f, part in enumerate(out._parts):
        value, bound = fp64.evaluate_with_bound(part.closure, (B, H, W),
                                                [data[id(p.id)] for p in fp64._param_order(part.closure)],
                                                mode="fast")
        assert fp64.within_bound(got[..., f], value, bound).all(), f
