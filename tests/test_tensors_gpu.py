"""The reference's known answers through the gfx950 back-end, and agreement with the
oracle (the reference's kernels run as host C) on the same graphs."""
import re

import numpy as np
import pytest

import spec_cases
from oracle import cpu_tensors, fp64

pytestmark = pytest.mark.gpu


def _flat(t):
    return t.flatArray()


@pytest.mark.parametrize("name", spec_cases.CASES)
def test_known_answers_on_gpu(name, gpu_tensors):
    spec_cases.check(name, gpu_tensors, lambda t: t.toString(), _flat)


def test_repeated_to_string(gpu_tensors):
    t = gpu_tensors.Tensor(42.0)
    for _ in range(1000):
        assert t.toString() == "42.0"


def test_kernel_cache_hit_for_equal_fill(gpu_tensors):
    """TensorsSpec.scala:37-55: a structurally equal closure finds the cached kernel."""
    T = gpu_tensors
    filled = T.Tensor.fill(42.0, (2, 3, 5))
    a = filled.flatArray()
    assert a.size == 30 and np.all(a == 42.0)
    assert T.kernel_cache_contains(filled.closure, filled.shape)
    zeros2 = T.Tensor.fill(42.0, (2, 3, 5))
    assert T.kernel_cache_contains(zeros2.closure, zeros2.shape)


def test_ragged_input_raises(gpu_tensors):
    from compute import IllegalArgumentException
    with pytest.raises(IllegalArgumentException):
        gpu_tensors.Tensor([[1.0], [3.0, 4.0]])


def test_do_cache_lifetime(gpu_tensors):
    """README.md:172-229: the cached buffer is reused until released."""
    T = gpu_tensors
    a = T.Tensor([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    b = T.Tensor([[7.0, 8.0, 9.0], [10.0, 11.0, 12.0]])
    c = T.Tensor([[13.0, 14.0, 15.0], [16.0, 17.0, 18.0]])
    result = a * b + c
    launches = T.launches
    cached = result.doCache()
    assert T.launches == launches + 1
    assert cached.toString() == "[[20.0,30.0,42.0],[56.0,72.0,90.0]]"
    assert cached.toString() == "[[20.0,30.0,42.0],[56.0,72.0,90.0]]"
    assert T.launches == launches + 1  # no recomputation
    tmp = T.Tensor.exp(cached)
    with np.errstate(over="ignore"):  # exp(90) is beyond FLT_MAX: +inf, as the kernel gives
        want = np.exp(np.array([[20, 30, 42], [56, 72, 90]], np.float64)).astype(np.float32)
    assert fp64.ulp_distance(tmp.flatArray(), want).max() <= 2
    cached.release()
    from compute import _capi as capi
    with pytest.raises(capi.InvalidMemObject):
        cached.toString()


@pytest.mark.parametrize("seed", range(8))
def test_random_graphs_agree_with_oracle(gpu_tensors, seed):
    """Random elementwise/view graphs with fusion barriers: bit-exact with the reference's
    kernels (host C, no FMA) for + - * / min max and views, <= 2 ulp with exp/tanh/log."""
    T = gpu_tensors
    rng = np.random.default_rng(seed)
    from oracle import inputs
    shape = tuple(int(s) for s in rng.integers(1, 9, size=rng.integers(1, 5)))
    n = int(np.prod(shape))
    x = T.Tensor(inputs.normal(n, 10 + seed).reshape(shape))
    y = T.Tensor(inputs.uniform(n, 20 + seed, 0.5, 2.0).reshape(shape))
    perm = list(rng.permutation(len(shape)))
    xp = x.permute(perm).permute(list(np.argsort(perm)))  # composed views
    off = [int(v) for v in rng.integers(-2, 3, size=len(shape))]
    xt = x.translate(off)
    e = (xp * y + xt) / y - T.Tensor.max(xt, x)
    if seed % 2:
        e = T.Tensor.tanh(e) + T.Tensor.exp(-T.Tensor.abs(e))
    if seed % 3 == 0:
        e = e.nonInline + x  # a fusion barrier
    got = e.flatArray()
    want = cpu_tensors.materialize(e)[:n]
    if seed % 2:
        # transcendentals: each within 2 ulp of its own inputs, propagated through the graph
        leaves = [cpu_tensors.materialize(p.id) for p in _params(e.closure)]
        value, bound = fp64.evaluate_with_bound(e.closure, e.shape, leaves, ulps=2.0)
        ok = fp64.within_bound(got, value, bound)
        assert ok.all(), (got[~ok][:4], value.ravel()[~ok][:4], bound.ravel()[~ok][:4])
    else:
        assert fp64.ulp_distance(got, want).max() == 0


def _params(closure):
    from compute import trees
    return trees.parameters(closure)


CONV_CASES = [
    # (depth, filters, vector width): 4 wide; constant table + a row of 3 read per element;
    # table + rows of 8 and 12 in 16-byte chunks; no table (18 + 2 constants), the 2 filters'
    # weights read as rows
    (3, 4, 4), (3, 4, 1), (8, 4, 1), (12, 2, 1), (1, 2, 1),
    (8, 4, 0),  # no launch shape given: a body this large runs one element per vector
]


@pytest.mark.parametrize("D,F,V", CONV_CASES)
def test_convolution_idiom_bit_exact(D, F, V):
    """The reference's convolution idiom (TensorsSpec.scala:144-210, benchmarks.scala:454-548:
    a join of per-filter sums of translated input channels times broadcast weights) on random
    data with padding at every edge: one kernel, bit-exact with the FP32 program (the sums
    keep the fold's order, no contraction). The same translated view taken once per term is
    gathered once. The weights and biases, when there are 32 or more, are read from a table in
    LDS, and at one element per vector a pixel's channels are read as one row: neither changes
    a bit."""
    from compute import Tensors, trees
    from oracle import inputs
    T = Tensors(0, vector_width=V)
    try:
        B, H, W = 2, 9, 12 if V in (0, 4) else 11
        x = inputs.normal(B * H * W * D, 61).reshape(B, H, W, D)
        w = inputs.normal(3 * 3 * D * F, 62).reshape(3, 3, D, F)
        b = inputs.normal(F, 63)
        tx, tw, tb = T.Tensor(x), T.Tensor(w), T.Tensor(b)
        out = spec_cases.convolute(T, tx, tw, tb)
        clo = trees.Concatenate([p.closure for p in out._parts])
        got = np.asarray(out.flatArray(), np.float32)
        data = {id(tx): x, id(tw): w, id(tb): b}
        leaves = [data[id(p.id)] for p in fp64._param_order(clo)]  # the kernel's first-visit order
        want = fp64.evaluate(clo, (B, H, W), leaves, "fp32").ravel()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        src = T.kernel_source(clo, (B, H, W))
        assert "nodes, 3 leaves" in src
        if V == 0:
            assert ", V=1 U=1" in src.split("\n", 1)[0]
            V = 1
        staged = 9 * D * F + F >= 32
        assert ("csmi_k[" in src) == staged
        # rows read together: a pixel's channels, and without the table the filters' weights
        assert bool(re.search(r"\br\d+_ok = live\[u\]", src)) == (V == 1 and (D > 1 or (not staged and F > 1)))
        assert bool(re.search(r"\br\d+_c0 = \*\(const csmi_f4\*\)", src)) == (V == 1 and D % 4 == 0)
    finally:
        T.close()
