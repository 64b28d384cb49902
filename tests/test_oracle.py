"""The oracle against the reference's known answers and against mathematics (CPU)."""
import numpy as np
import pytest

import oracle
import spec_cases
from oracle import cpu_tensors, fp64, inputs, refcl


def _ct_string(t):
    return cpu_tensors.to_string(t)


def _ct_flat(t):
    n = 1
    for s in t.shape:
        n *= s
    return cpu_tensors.materialize(t)[:n]


@pytest.mark.parametrize("name", spec_cases.CASES)
def test_refcl_known_answers(name, offline_tensors):
    """refcl-host reproduces every known answer of the reference's tests and docs."""
    spec_cases.check(name, offline_tensors, _ct_string, _ct_flat)


def test_random_bit_exact():
    """TensorsSpec.scala:402-409, and the numpy hash agrees with the C restatement."""
    r = oracle.random(9, 12345)
    assert cpu_tensors.refcl_string(r, (3, 3)) == spec_cases.GOLDEN["random"]["expect"]
    h = inputs.wang_hash(np.arange(9, dtype=np.uint32) ^ np.uint32(12345))
    assert np.array_equal((h.astype(np.float32) / np.float32(4294967296.0)), r)


def test_random_normal_known_values():
    exp = spec_cases.GOLDEN["random_normal"]["expect"]
    got = cpu_tensors.refcl_string(oracle.random_normal(27, 54321), (3, 3, 3))
    assert got == exp


def test_sequential_sum_matches_fp64_bound():
    x = inputs.normal(100003, 7)
    s = oracle.sequential_sum(x)
    exact = np.sum(x.astype(np.float64))
    assert abs(float(s) - exact) <= 100003 * fp64.EPS32 * np.sum(np.abs(x.astype(np.float64)))
    assert oracle.sequential_sum(np.full(64, 15625.0, np.float32)) == np.float32(1e6)
    assert oracle.sequential_sum(np.arange(1, 17, dtype=np.float32)) == np.float32(136)
    assert oracle.sequential_sum(np.zeros(0, np.float32)) == 0


def test_parallel_sum_power_of_two_and_quirk():
    """The reference's GPU tree is exact for power-of-two local sizes and drops the odd
    element otherwise (Tensors.scala:375-387)."""
    x = np.ones(256 * 4, np.float32)
    assert oracle.parallel_sum(x, 256, 4) == np.float32(1024)
    # length 6 < work group: local size 6, offsets 3 -> stops, so lane 2 is never added
    y = np.array([1, 2, 4, 8, 16, 32], np.float32)
    assert oracle.parallel_sum(y, 256, 4) == np.float32(1 + 2 + 8 + 16)


def test_fp64_evaluator_closed_forms(offline_tensors):
    T = offline_tensors
    x = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    t = T.Tensor(x)
    p = t.permute([2, 0, 1])
    assert np.array_equal(fp64.evaluate(p.closure, p.shape, [x]), np.transpose(x, (2, 0, 1)))
    tr = t.translate([0, 1, -2])
    exp = np.zeros_like(x)
    exp[:, 1:, :2] = x[:, :2, 2:]
    assert np.array_equal(fp64.evaluate(tr.closure, tr.shape, [x]), exp)
    b = T.Tensor(x[:, :, 0]).broadcast((2, 3, 5))
    assert np.array_equal(fp64.evaluate(b.closure, b.shape, [x[:, :, 0].copy()]),
                          np.repeat(x[:, :, :1], 5, axis=2))


def _random_closure(T, rng, leaves, depth):
    if depth == 0 or rng.random() < 0.25:
        return leaves[rng.integers(len(leaves))]
    op = rng.choice(["+", "-", "*", "/", "max", "min", "exp", "tanh", "neg", "abs"])
    if op in ("exp", "tanh"):
        return getattr(T.Tensor, op)(_random_closure(T, rng, leaves, depth - 1) * T.Tensor.fill(0.25, leaves[0].shape))
    if op == "neg":
        return -_random_closure(T, rng, leaves, depth - 1)
    if op == "abs":
        return T.Tensor.abs(_random_closure(T, rng, leaves, depth - 1))
    a = _random_closure(T, rng, leaves, depth - 1)
    b = _random_closure(T, rng, leaves, depth - 1)
    if op in ("max", "min"):
        return getattr(T.Tensor, op)(a, b)
    return {"+": a.__add__, "-": a.__sub__, "*": a.__mul__, "/": a.__truediv__}[op](b)


@pytest.mark.parametrize("seed", range(6))
def test_refcl_agrees_with_fp32_evaluator(seed, offline_tensors):
    """Two independent restatements of the kernel semantics agree on random DAGs with views:
    bit-exact for + - * / min max, <= 2 ulp where exp/tanh appear."""
    T = offline_tensors
    rng = np.random.default_rng(seed)
    x = inputs.normal(4 * 5 * 6, 100 + seed).reshape(4, 5, 6)
    y = inputs.uniform(5 * 4 * 6, 200 + seed).reshape(5, 4, 6)
    tx, ty = T.Tensor(x), T.Tensor(y)
    views = [tx, ty.permute([1, 0, 2]), tx.translate([1, 0, -2]), T.Tensor(x[:, :1, :].copy()).broadcast((4, 5, 6)),
             T.Tensor(x[:, 0, 0].copy()).broadcast((4, 5, 6))]
    expr = _random_closure(T, rng, views, 4)
    if not hasattr(expr, "closure") or expr.shape != (4, 5, 6):
        return
    k = refcl.Kernel(expr.closure, expr.shape)
    leaf_data = [cpu_tensors.materialize(p.id) for p in k.leaves]
    got = k(leaf_data)
    want = fp64.evaluate(expr.closure, expr.shape, leaf_data, "fp32").ravel()
    assert fp64.ulp_distance(got, want).max() <= 2


def test_refcl_build_failures_match_reference_quirks(offline_tensors):
    """Float `%` and infinite literals do not build in the reference's kernel text."""
    T = offline_tensors
    a = T.Tensor([1.0, 2.0])
    with pytest.raises(refcl.BuildProgramFailure):
        refcl.Kernel((a % a).closure, (2,))
    with pytest.raises(refcl.BuildProgramFailure):
        refcl.Kernel(T.Tensor.fill(float("inf"), (2,)).closure, (2,))


def test_to_literal_decimal_format():
    assert refcl.to_literal(2.0 / 3.0) == "0.667"
    assert refcl.to_literal(0.0625) == "0.062"
    assert refcl.to_literal(-1.0) == "-1"
    assert refcl.to_literal(0.5) == "0.5"
    assert refcl.to_literal(3.0) == "3"
    assert refcl.java_float_to_string(1e-5) == "1.0E-5"
    assert refcl.java_float_to_string(1e7) == "1.0E7"
    assert refcl.java_float_to_string(123456.7) == "123456.7"


def test_gemm_oracles_agree():
    a = inputs.uniform(64 * 48, 1).reshape(64, 48)
    b = inputs.uniform(48 * 40, 2).reshape(48, 40)
    c = oracle.gemm_seqk(a, b)
    assert fp64.gemm_normalized_error(c, a, b) < 48 * fp64.EPS32
    small = np.array([[1, 2, 3], [4, 5, 6]], np.float32)
    big = np.array([[7, 8, 9, 10], [11, 12, 13, 14], [15, 16, 17, 18]], np.float32)
    assert oracle.gemm_seqk(small, big).tolist() == [[74, 80, 86, 92], [173, 188, 203, 218]]


def test_row_fold_and_fp64_stats():
    a = inputs.heavy_tails(inputs.normal(64 * 100, 5), 6).reshape(64, 100)
    s, m = oracle.row_fold(a)
    s64, m64, sa, total = oracle.row_stats64(a)
    assert np.array_equal(m, m64)
    assert np.array_equal(m, a.max(axis=1))
    assert np.all(np.abs(s - s64) <= 100 * fp64.EPS32 * sa)
    assert abs(total - a.astype(np.float64).sum()) <= 1e-9 * np.abs(a).sum()


@pytest.mark.parametrize("seed", range(4))
def test_error_bound_holds_for_host_libm_and_catches_planted_errors(seed, offline_tensors):
    """The propagated bound admits the host libm's results (each op within 1 ulp) and
    rejects results with a planted 6-ulp error in one transcendental."""
    T = offline_tensors
    rng = np.random.default_rng(seed)
    x = inputs.normal(6 * 7, 300 + seed).reshape(6, 7)
    y = inputs.uniform(6 * 7, 400 + seed, 0.5, 2.0).reshape(6, 7)
    tx, ty = T.Tensor(x), T.Tensor(y)
    e = T.Tensor.tanh(tx * ty) + T.Tensor.exp(-T.Tensor.abs(tx)) - T.Tensor.log(ty) * T.Tensor.sqrt(ty)
    k = refcl.Kernel(e.closure, e.shape)
    got = k([x, y])
    value, bound = fp64.evaluate_with_bound(e.closure, e.shape, [x, y], ulps=2.0)
    assert fp64.within_bound(got, value, bound).all()
    # plant: the exp term off by 6 ulp
    ex = np.exp(-np.abs(x.astype(np.float64))).astype(np.float32)
    bad_ex = ex
    for _ in range(6):
        bad_ex = np.nextafter(bad_ex, np.float32(np.inf))
    th = np.tanh((x * y).astype(np.float64)).astype(np.float32)
    lg = np.log(y.astype(np.float64)).astype(np.float32)
    sq = np.sqrt(y.astype(np.float64)).astype(np.float32)
    planted = (th + bad_ex) - lg * sq
    ok = fp64.within_bound(planted, value, bound)
    assert not ok.all()


def test_fast_kernels_leave_the_fp_environment_alone():
    """Loading a refcl-host kernel built in the fast mode must not switch the process to
    flush-to-zero / denormals-are-zero (gcc links crtfastmath.o into shared objects built with
    -funsafe-math-optimizations): the FP64 oracle in the same process needs subnormals."""
    import numpy as np
    from compute import Tensors
    from oracle import refcl
    T = Tensors(offline=True)
    x = np.array([1e-41, 2.0], np.float32)
    e = T.Tensor(x) * T.Tensor(x)
    refcl.Kernel(e.closure, e.shape, "fast")([x, x])
    tiny = np.array([1e-41], np.float32)
    assert np.sqrt(tiny)[0] > 0 and (tiny * np.float32(2.0))[0] > 0
