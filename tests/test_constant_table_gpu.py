"""Constant tables (single elements of a small leaf read through LDS) where workgroups carry
idle lanes or add their elements instead of storing them: a weighted sum of a transposed leaf
in the tiled launch order (padded tiles) is bit-identical to the linear order and to the
element-wise program, and its fused sum is within the compensated sum's bound."""
import functools

import numpy as np
import pytest

from compute import _capi as capi
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu


def weighted(T, x, w, shift):
    """sum_k w[k] * translate(x^T, k - shift): 40 constants of w, each broadcast."""
    xt = T.Tensor(x).permute([1, 0])
    ws = list(T.Tensor(w).split(0))
    terms = [xt.translate([0, k - shift]) * ws[k].broadcast(xt.shape) for k in range(len(ws))]
    return functools.reduce(lambda a, b: a + b, terms)


@pytest.mark.parametrize("shape", [(318, 934), (96, 64), (33, 70)])
def test_table_in_tiled_order(shape):
    from compute import Tensors
    x = inputs.uniform(shape[0] * shape[1], 911).reshape(shape)
    w = inputs.uniform(40, 912)
    got, src = {}, {}
    for order in (capi.ORDER_AUTO, capi.ORDER_LINEAR):
        with Tensors(0, vector_width=1, launch_order=order) as T:
            e = weighted(T, x, w, 20)
            src[order] = T.kernel_source(e.closure, e.shape)
            got[order] = e.flatArray()
    assert "csmi_k[" in src[capi.ORDER_AUTO] and "csmi_k[" in src[capi.ORDER_LINEAR]
    print("shape %s: tiled %s" % (shape, "dt * " in src[capi.ORDER_AUTO]), flush=True)
    assert np.array_equal(got[capi.ORDER_AUTO].view(np.uint32), got[capi.ORDER_LINEAR].view(np.uint32))
    # the FP32 program (the oracle's evaluation in FP32, the fold's order, no contraction)
    with Tensors(offline=True, vector_width=1) as U:
        e = weighted(U, x, w, 20)
        data = {id(leaf): arr for leaf, arr in zip(_leaves(e), (x, w))}
        leaves = [data[id(p.id)] for p in fp64._param_order(e.closure)]
        want = fp64.evaluate(e.closure, e.shape, leaves, "fp32").ravel()
    assert np.array_equal(got[capi.ORDER_AUTO].view(np.uint32), want.view(np.uint32))


def _leaves(e):
    """The two leaf tensors of `weighted`, x first."""
    ids = {}
    for p in fp64._param_order(e.closure):
        ids.setdefault(id(p.id), p.id)
    ts = list(ids.values())
    return sorted(ts, key=lambda t: -int(np.prod(t.shape)))


def test_table_in_a_fused_sum():
    from compute import Tensors
    x = inputs.uniform(300 * 500, 913).reshape(300, 500)
    w = inputs.uniform(40, 914)
    with Tensors(0) as T:
        e = weighted(T, x, w, 7)
        vals = e.flatArray().astype(np.float64)
        before = T.launches
        s = e.sum().flatArray()[0]
        assert T.launches - before == 1
        sum_src = T.kernel_source(e.closure, e.shape)
    assert "csmi_k[" in sum_src
    assert abs(float(s) - vals.sum()) <= 4 * fp64.EPS32 * np.abs(vals).sum()
