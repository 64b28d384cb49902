"""The cost-model network on the device: gradients against the FP64 oracle, and training."""
import ctypes

import numpy as np
import pytest

from compute import costmodel
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _upload(T, a):
    from gpu_util import upload
    return upload(T, a)


@pytest.mark.parametrize("widths", [[23, 64, 64, 64, 1], [5, 7, 1], [3, 1]])
def test_forward_and_gradients_match_fp64(gpu_tensors, widths):
    T = gpu_tensors
    net = costmodel.Network(T, widths, seed=3)
    rng = np.random.default_rng(1)
    n = 300
    X = rng.standard_normal((n, widths[0])).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    p = net.get_params()
    xb, tb = _upload(T, X), _upload(T, t)
    yb = T._alloc(n * 4)
    net.forward_device(xb.ptr, n, yb.ptr)
    y = yb.to_numpy(n)
    y64 = mlp_ref.forward(p, widths, X)[-1][:, 0]
    assert np.max(np.abs(y - y64)) <= 1e-5 * max(1.0, np.max(np.abs(y64)))
    loss_buf, g_buf = net.gradients_device(xb.ptr, tb.ptr, n)
    g = g_buf.to_numpy(net.count)
    loss = float(loss_buf.to_numpy(1)[0])
    loss64, g64 = mlp_ref.loss_and_grad(p, widths, X, t)
    assert abs(loss - loss64) <= 1e-5 * max(1.0, loss64)
    scale = np.max(np.abs(g64))
    err = np.max(np.abs(g - g64)) / scale
    print("widths %s: max gradient error %.3g of the largest gradient" % (widths, err))
    assert err <= 1e-4
    for b in (xb, tb, yb):
        b.release()
    net.release()


def test_training_reduces_loss_and_fits_a_known_function(gpu_tensors):
    T = gpu_tensors
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(4096, 6)).astype(np.float32)
    us = np.exp(1.5 * X[:, 0] - X[:, 1] * X[:, 2] + 0.5 * np.abs(X[:, 3])) * 10.0
    m = costmodel.CostModel(T, 6, hidden=32, layers=2, seed=4)
    hist = m.fit(X, us, epochs=60, batch=256, lr=3e-3)
    assert hist[-1] < 0.25 * hist[0], hist[::10]
    pred = m.predict_log_us(X[:512])
    resid = np.std(pred - np.log(us[:512]))
    print("fit residual std %.3f (target std %.3f)" % (resid, np.std(np.log(us))))
    assert resid < 0.35 * np.std(np.log(us))
    m.release()
