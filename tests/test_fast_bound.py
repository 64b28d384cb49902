"""The FP64 oracle's bound for the fast math mode, pinned on the CPU against simulated fast
implementations: FMA contraction of a*b+c (exact product, one rounding), a reassociated
three-term sum, exp as 2^(x*log2 e) with the product rounded in FP32, and a division as a
product with the rounded reciprocal. Each must sit inside the fast bound, and the contracted
a*b+c must fall outside the precise bound somewhere (the fast model is not vacuous)."""
from fractions import Fraction

import numpy as np

from compute import Tensors
from oracle import fp64, inputs


def _tensors():
    return Tensors(offline=True)


def _fma32(a, b, c):
    """Correctly rounded a*b+c in FP32, exactly (rational arithmetic on a small sample)."""
    out = np.empty(a.size, np.float32)
    for i, (x, y, z) in enumerate(zip(a.ravel().tolist(), b.ravel().tolist(), c.ravel().tolist())):
        exact = Fraction(x) * Fraction(y) + Fraction(z)
        f = np.float32(float(exact))  # float() rounds to FP64; FP64 -> FP32 may double-round
        # settle double rounding: pick the FP32 neighbour nearest the exact value
        cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        out[i] = min(cands, key=lambda v: abs(Fraction(float(v)) - exact))
    return out


def test_contracted_fma_is_inside_the_fast_bound():
    T = _tensors()
    n = 4096
    a = inputs.uniform(n, 11).reshape(64, 64)
    b = inputs.uniform(n, 12).reshape(64, 64)
    c = inputs.uniform(n, 13).reshape(64, 64)
    # near cancellation for a quarter of the entries: c = -RN(a*b) + tiny
    c.ravel()[::4] = (-(a * b) + np.float32(1e-7) * inputs.uniform(n, 14).reshape(64, 64)).ravel()[::4]
    e = T.Tensor(a) * T.Tensor(b) + T.Tensor(c)
    got = _fma32(a, b, c)
    value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [a, b, c], mode="fast")
    assert fp64.within_bound(got, value, bound).all()
    value_p, bound_p = fp64.evaluate_with_bound(e.closure, e.shape, [a, b, c])
    assert not fp64.within_bound(got, value_p, bound_p).all()


def test_reassociated_sum_is_inside_the_fast_bound():
    T = _tensors()
    n = 1 << 14
    x, y, z = (inputs.normal(n, s) * np.float32(10.0 ** (k - 1)) for k, s in enumerate((21, 22, 23)))
    e = T.Tensor(x) + T.Tensor(y) + T.Tensor(z)
    other_order = x + (y + z)  # FP32 throughout
    value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [x, y, z], mode="fast")
    assert fp64.within_bound(other_order, value, bound).all()


def test_fast_exp_and_division_are_inside_the_fast_bound():
    T = _tensors()
    n = 1 << 16
    x = inputs.uniform(n, 31, -87.0, 88.0)
    e = T.Tensor.exp(T.Tensor(x))
    l2e = np.float32(1.4426950408889634)
    t = (x * l2e).astype(np.float32)  # the rounded product
    # the hardware 2^t within one ulp: take the far neighbour of the correctly rounded value
    fast = np.exp2(t.astype(np.float64)).astype(np.float32)
    fast = np.nextafter(fast, np.float32(np.inf))
    value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [x], mode="fast")
    assert fp64.within_bound(fast, value, bound).all()
    # the precise bound does not hold for it at large |x|
    vp, bp = fp64.evaluate_with_bound(e.closure, e.shape, [x])
    assert not fp64.within_bound(fast, vp, bp).all()
    p = inputs.uniform(n, 32, 0.5, 2.0)
    q = inputs.uniform(n, 33, 0.5, 2.0)
    d = T.Tensor(p) / T.Tensor(q)
    recip = (np.float32(1.0) / q).astype(np.float32)
    by_recip = (p * recip).astype(np.float32)
    value, bound = fp64.evaluate_with_bound(d.closure, d.shape, [p, q], mode="fast")
    assert fp64.within_bound(by_recip, value, bound).all()


def test_bound_through_an_overflowing_exp():
    """exp(100) overflows to +inf in FP32 and tanh(inf) is exactly 1: the propagated bound is
    a number (infinite where the derivative overflows), never NaN from 0 * inf, in both
    modes; an exact input adds no error to an op whose derivative overflows."""
    T = _tensors()
    x = np.array([100.0, 1.0, -100.0], np.float32)
    e = T.Tensor.tanh(T.Tensor.exp(T.Tensor(x)))
    for mode in ("precise", "fast"):
        value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [x], mode=mode)
        assert not np.isnan(bound).any(), mode
        assert value[0] == 1.0 and bound[1] < 1e-5
    z = T.Tensor.exp(T.Tensor(x)) * T.Tensor(np.zeros(3, np.float32))
    value, bound = fp64.evaluate_with_bound(z.closure, z.shape, [x, np.zeros(3, np.float32)])
    assert np.isnan(value[0]) and not np.isnan(bound[1:]).any()
