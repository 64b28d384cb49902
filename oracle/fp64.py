"""numpy oracle for fused expressions, reductions and GEMM (test infrastructure).

`evaluate(closure, shape, leaves, mode)` interprets a closure DAG (the facade's node
objects, duck-typed by class name) over the output coordinates with numpy:

  mode "fp64":  every operation in FP64 on the FP32 inputs;
  mode "fp32":  every operation's result rounded to FP32 (+ - * / fmod min max exactly
                as IEEE FP32; exp, log, tanh, sqrt computed in FP64 and rounded once),
                i.e. the FP32 program with correctly rounded operations.

The view semantics are the reference's (OpenCLKernelBuilder.scala:348-411, 435-455):
`idx = (int)(sum of coordinate * coefficient + constant)`, truncated toward zero, with
each coefficient rounded to three decimals as the kernel text prints it
(OpenCLKernelBuilder.scala:14-32); a view's source index outside [0, dim) reads the
padding; an all-zero row reads index 0 unchecked; a bare leaf checks only idx < dim.
View chains compose as homogeneous matrix products (the algebra of
NDimensionalAffineTransform.scala:28-95), written independently of the product code.
"""
import decimal

import numpy as np


def _round3(v):
    return float(decimal.Decimal(float(v)).quantize(decimal.Decimal("0.001"), rounding=decimal.ROUND_HALF_EVEN))


def _homogeneous(matrix, rows, cols_in):
    m = np.zeros((rows + 1, cols_in + 1))
    m[:rows, :] = np.asarray(matrix, dtype=np.float64).reshape(rows, cols_in + 1)
    m[rows, cols_in] = 1.0
    return m


def _resolve(node):
    """(kind, payload): ("fill", float node) or ("array", param, matrix or None)."""
    kind = type(node).__name__
    if kind == "ArrayParameter":
        return ("array", node, None)
    if kind == "Fill":
        return ("fill", node.element)
    if kind == "Transform":
        inner = _resolve(node.array)
        if inner[0] == "fill":
            return inner
        _, param, m = inner
        src_rank = len(param.shape)
        view_in = len(node.matrix) // (node.rank + 1)  # rows = rank of the inner space
        mine = _homogeneous(node.matrix, view_in, node.rank)
        if m is None:
            return ("array", param, mine)
        return ("array", param, m @ mine)  # inner map after this view
    raise TypeError("not an array node: " + kind)


def _gather(param, matrix, data, out_shape):
    src_shape = tuple(param.shape)
    data = np.asarray(data, dtype=np.float32).reshape(src_shape) if src_shape else np.asarray(data, np.float32).ravel()
    pad = np.float32(param.padding)
    r = len(out_shape)
    coords = np.indices(out_shape, dtype=np.int64) if r else np.zeros((0,), np.int64)
    if len(src_shape) == 0:
        return np.full(out_shape, data.ravel()[0], dtype=np.float32)
    ok = np.ones(out_shape, dtype=bool)
    idx = []
    for y in range(len(src_shape)):
        if matrix is None:
            if y >= r:
                raise ValueError("leaf rank exceeds the kernel rank")
            iy = coords[y]
            ok &= iy < src_shape[y]
        else:
            row = matrix[y]
            coefs = [(x, _round3(row[x])) for x in range(r) if row[x] != 0.0]
            const = _round3(row[r]) if row[r] != 0.0 else None
            if not coefs and const is None:
                iy = np.zeros(out_shape, dtype=np.int64)
            else:
                if all(float(c).is_integer() for _, c in coefs) and (const is None or float(const).is_integer()):
                    acc = np.zeros(out_shape, dtype=np.int64)
                    for x, c in coefs:
                        acc = acc + coords[x] * int(c)
                    if const is not None:
                        acc = acc + int(const)
                    iy = acc
                else:
                    acc = np.zeros(out_shape, dtype=np.float64)
                    for x, c in coefs:
                        acc = acc + coords[x] * c
                    if const is not None:
                        acc = acc + const
                    iy = np.trunc(acc).astype(np.int64)
                ok &= (iy >= 0) & (iy < src_shape[y])
        idx.append(iy)
    safe = [np.where(ok, iy, 0) for iy in idx]
    vals = data[tuple(safe)]
    return np.where(ok, vals, pad).astype(np.float32)


def evaluate(closure, shape, leaves, mode="fp64"):
    """Values of `closure` at every coordinate of `shape` (tuples: last axis = element).
    `leaves` maps each ArrayParameter (by identity) or leaf index to its data."""
    shape = tuple(int(s) for s in shape)
    f32 = mode == "fp32"
    dt = np.float32 if f32 else np.float64
    memo = {}
    params = _param_order(closure)
    data_of = {}
    for i, p in enumerate(params):
        data_of[id(p)] = leaves[i] if isinstance(leaves, (list, tuple)) else leaves[id(p)]

    def ev(node):
        key = id(node)
        if key in memo:
            return memo[key]
        kind = type(node).__name__
        if kind == "FloatLiteral":
            v = np.full(shape, node.value, dtype=np.float32).astype(dt)
        elif kind == "Extract":
            res = _resolve(node.array)
            if res[0] == "fill":
                v = ev(res[1])
            else:
                _, param, m = res
                v = _gather(param, m, data_of[id(param)], shape).astype(dt)
        elif kind == "Unary":
            a = ev(node.operand)
            with np.errstate(all="ignore"):
                if node.op == "neg":
                    v = -a
                elif node.op == "abs":
                    v = np.abs(a)
                else:
                    fn = {"sqrt": np.sqrt, "tanh": np.tanh, "exp": np.exp, "log": np.log}[node.op]
                    v = fn(a.astype(np.float64)).astype(dt)
        elif kind == "Binary":
            a, b = ev(node.lhs), ev(node.rhs)
            with np.errstate(all="ignore"):
                v = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide, "%": np.fmod,
                     "min": np.fmin, "max": np.fmax}[node.op](a, b)
        else:
            raise TypeError("unexpected node " + kind)
        v = np.asarray(v, dtype=dt)
        memo[key] = v
        return v

    if type(closure).__name__ == "Concatenate":
        return np.stack([ev(e) for e in closure.elements], axis=-1)
    return ev(closure)


def _ulp32(v):
    """ulp of FP32 values (as FP64), ulp(0) = smallest subnormal; +inf for infinite or NaN
    values (and magnitudes beyond FP32), whose neighbourhood has no finite width."""
    v = np.abs(np.asarray(v, dtype=np.float64))
    with np.errstate(all="ignore"):
        f = v.astype(np.float32)
        nxt = np.nextafter(f, np.float32(np.inf))
        u = nxt.astype(np.float64) - f.astype(np.float64)
    return np.where(np.isfinite(u), u, np.inf)


# Error model of the fast math mode (the counterpart of -cl-unsafe-math-optimizations),
# stated as bounds that the device measurements over every FP32 input sit inside:
#   exp(x) = 2^(x * log2(e)) with the product rounded once: the rounding of the product
#            (relative 2^-24, and log2(e) itself rounded) scales with |x|, so the result is
#            within FAST_EXP_ULPS[0] + FAST_EXP_ULPS[1] * |x| ulp (the rounded product alone can
#            cost 1.25 |x| ulp; measured on the device: at most 1 ulp
#            for |x| < 1, 59 for |x| < 64, 64 up to the overflow at 88.7); results below
#            2^-126 may be flushed to zero (FAST_FLUSH absolute);
#   log(x) from the hardware base-2 logarithm: within FAST_LOG_ULPS ulp everywhere
#            (measured 2, subnormal arguments included);
#   sqrt from the hardware square root: FAST_SQRT_ULPS ulp, and a subnormal argument may
#            be flushed to a zero of its sign first (an error below 2^-63, and -0 rather
#            than NaN for a negative one);
#   tanh keeps the precise mode's function (2 ulp).
FAST_EXP_ULPS = (2.0, 1.5)
FAST_FLUSH = 2.0 ** -126
FAST_LOG_ABS = 0.0
FAST_LOG_ULPS = 2.0
FAST_SQRT_ULPS = 2.0


def _times(x, y):
    """x * y for propagated bounds, where a zero factor gives 0 even against an infinite one
    (an exact input of an operation whose derivative overflows adds no error)."""
    with np.errstate(all="ignore"):
        return np.where((np.asarray(x) == 0) | (np.asarray(y) == 0), 0.0, np.asarray(x, np.float64) * y)


def evaluate_with_bound(closure, shape, leaves, ulps=2.0, mode="precise"):
    """(value, bound): `value` is the FP32 program with correctly rounded operations
    (mode "fp32" of `evaluate`); `bound` is a first-order bound on how far an
    implementation may land from it when + - * / min max are correctly rounded and each of
    exp, tanh, log and sqrt is within `ulps` ulp of the exact result of its own inputs.
    Errors propagate through each operation by its derivative (mean value bounds).

    mode "fast" bounds the fast math mode instead, where the compiler may contract a
    product into an addition (the product is then not rounded), reassociate sums and
    divide by multiplying with a reciprocal: every + - * / contributes one ulp of the
    magnitude it works at (for a chain of additions, the sum of the magnitudes of its
    terms, which bounds every partial sum of any association), a division two, and exp,
    log and sqrt follow the FAST_* model above."""
    fast = mode == "fast"
    shape = tuple(int(s) for s in shape)
    params = _param_order(closure)
    data_of = {id(p): (leaves[i] if isinstance(leaves, (list, tuple)) else leaves[id(p)])
               for i, p in enumerate(params)}
    memo = {}
    mags = {}  # fast mode: sum of the term magnitudes of an addition chain

    def mag_of(node, v, b):
        if id(node) in mags:
            return mags[id(node)]
        return np.abs(v.astype(np.float64)) + b

    def ev(node):
        key = id(node)
        if key in memo:
            return memo[key]
        kind = type(node).__name__
        zero = np.zeros(shape)
        if kind == "FloatLiteral":
            out = (np.full(shape, node.value, dtype=np.float32), zero)
        elif kind == "Extract":
            res = _resolve(node.array)
            if res[0] == "fill":
                out = ev(res[1])
            else:
                _, param, m = res
                out = (_gather(param, m, data_of[id(param)], shape), zero)
        elif kind == "Unary":
            a, ba = ev(node.operand)
            a64 = a.astype(np.float64)
            with np.errstate(all="ignore"):
                if node.op == "neg":
                    out = (-a, ba)
                elif node.op == "abs":
                    out = (np.abs(a), ba)
                else:
                    v = {"sqrt": np.sqrt, "tanh": np.tanh, "exp": np.exp, "log": np.log}[node.op](a64).astype(np.float32)
                    if node.op == "exp":
                        deriv = np.exp(a64 + ba)
                    elif node.op == "tanh":
                        deriv = np.ones(shape)
                    elif node.op == "log":
                        deriv = 1.0 / np.maximum(np.abs(a64) - ba, 1e-300)
                    else:
                        deriv = 0.5 / np.sqrt(np.maximum(np.abs(a64) - ba, 1e-300))
                    own = (ulps + 0.5) * _ulp32(v)
                    if fast and node.op == "exp":
                        own = (FAST_EXP_ULPS[0] + FAST_EXP_ULPS[1] * (np.abs(a64) + ba) + 0.5) * _ulp32(v)
                        own = own + FAST_FLUSH
                    elif fast and node.op == "log":
                        own = FAST_LOG_ABS + (FAST_LOG_ULPS + 0.5) * _ulp32(v)
                    elif fast and node.op == "sqrt":
                        own = (FAST_SQRT_ULPS + 0.5) * _ulp32(v)
                        own = own + np.where(np.abs(a64) < FAST_FLUSH, 2.0 ** -63, 0.0)
                    out = (v, _times(deriv, ba) + own)
        elif kind == "Binary":
            a, ba = ev(node.lhs)
            b, bb = ev(node.rhs)
            a64, b64 = a.astype(np.float64), b.astype(np.float64)
            with np.errstate(all="ignore"):
                v = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide, "%": np.fmod,
                     "min": np.fmin, "max": np.fmax}[node.op](a, b)
                if node.op in ("+", "-"):
                    prop = ba + bb
                elif node.op == "*":
                    prop = _times(np.abs(a64), bb) + _times(np.abs(b64), ba) + _times(ba, bb)
                elif node.op == "/":
                    prop = (ba + _times(np.abs(a64 / b64), bb)) / np.maximum(np.abs(b64) - bb, 1e-300)
                elif node.op in ("min", "max"):
                    prop = np.maximum(ba, bb)
                else:
                    prop = np.where((ba > 0) | (bb > 0), np.inf, 0.0)
                rounding = np.where(prop > 0, _ulp32(np.abs(v.astype(np.float64)) + prop), 0.0)
                if fast and node.op in ("+", "-"):
                    m = mag_of(node.lhs, a, ba) + mag_of(node.rhs, b, bb)
                    mags[key] = m
                    rounding = _ulp32(m + prop)
                elif fast and node.op == "*":
                    rounding = _ulp32(np.abs(v.astype(np.float64)) + prop)
                elif fast and node.op == "/":
                    rounding = 2.0 * _ulp32(np.abs(v.astype(np.float64)) + prop)
                if node.op in ("min", "max"):
                    rounding = 0.0
                out = (v, prop + rounding)
        else:
            raise TypeError("unexpected node " + kind)
        memo[key] = out
        return out

    if type(closure).__name__ == "Concatenate":
        vals = [ev(e) for e in closure.elements]
        return np.stack([v for v, _ in vals], axis=-1), np.stack([b for _, b in vals], axis=-1)
    return ev(closure)


def within_bound(got, value, bound):
    """True where `got` is within `bound` of `value` (NaN matches NaN, inf matches inf)."""
    got = np.asarray(got, dtype=np.float32).ravel()
    value = np.asarray(value, dtype=np.float32).ravel()
    bound = np.asarray(bound, dtype=np.float64).ravel()
    with np.errstate(all="ignore"):
        diff = np.abs(got.astype(np.float64) - value.astype(np.float64))
    same = (got == value) | (np.isnan(got) & np.isnan(value))
    return same | (diff <= bound)


def _param_order(closure):
    seen, out, stack = set(), [], [closure]
    while stack:
        n = stack.pop()
        if id(n) in seen:
            continue
        seen.add(id(n))
        k = type(n).__name__
        if k == "ArrayParameter":
            out.append(n)
            continue
        kids = {"Transform": lambda: [n.array], "Extract": lambda: [n.array], "Fill": lambda: [n.element],
                "Unary": lambda: [n.operand], "Binary": lambda: [n.lhs, n.rhs],
                "Concatenate": lambda: list(n.elements)}.get(k, lambda: [])()
        stack.extend(reversed(kids))
    return out


# ----------------------------------------------------------------------------- metrics
def ulp_distance(a, b):
    """|a - b| in units in the last place of FP32, elementwise (NaN vs NaN = 0)."""
    a = np.asarray(a, dtype=np.float32).ravel()
    b = np.asarray(b, dtype=np.float32).ravel()
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    oa = np.where(ia < 0, -0x80000000 - ia, ia)
    ob = np.where(ib < 0, -0x80000000 - ib, ib)
    d = np.abs(oa - ob)
    both_nan = np.isnan(a) & np.isnan(b)
    d[both_nan] = 0
    one_nan = np.isnan(a) ^ np.isnan(b)
    d[one_nan] = np.iinfo(np.int64).max
    return d


EPS32 = 2.0 ** -23


def sum_bound(x, axis=None):
    """The reduction tolerance 4 * eps32 * sum |x| (FP64)."""
    return 4.0 * EPS32 * np.sum(np.abs(np.asarray(x, dtype=np.float64)), axis=axis)


def gemm_fp64(a, b):
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def gemm_normalized_error(c, a, b, c64=None):
    """max_ij |C - C64| / sum_k |a_ik| |b_kj|."""
    if c64 is None:
        c64 = gemm_fp64(a, b)
    denom = np.abs(np.asarray(a, np.float64)) @ np.abs(np.asarray(b, np.float64))
    denom[denom == 0] = 1.0
    return float(np.max(np.abs(np.asarray(c, np.float64) - c64) / denom))
