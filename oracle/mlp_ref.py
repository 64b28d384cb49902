"""FP64 restatement of the cost-model network (test infrastructure).

Forward and analytic backward of widths[0] -> ... -> widths[-1] with ReLU after every layer
but the last and a mean-squared-error loss, on the same flat parameter layout as the device
network ([W1 (in x out, row-major), b1, W2, b2, ...]). Pinned by central finite differences
(tests/test_mlp_oracle.py) and by closed forms on one-layer networks.
"""
import numpy as np


def unflatten(params, widths):
    params = np.asarray(params, dtype=np.float64)
    layers, off = [], 0
    for din, dout in zip(widths[:-1], widths[1:]):
        W = params[off:off + din * dout].reshape(din, dout)
        off += din * dout
        b = params[off:off + dout]
        off += dout
        layers.append((W, b))
    assert off == params.size
    return layers


def forward(params, widths, X):
    acts = [np.asarray(X, dtype=np.float64)]
    layers = unflatten(params, widths)
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ W + b
        acts.append(np.maximum(z, 0.0) if l < len(layers) - 1 else z)
    return acts


def loss_and_grad(params, widths, X, target):
    """(loss, flat gradient) of mean((net(X) - target)^2)."""
    acts = forward(params, widths, X)
    layers = unflatten(params, widths)
    n = acts[0].shape[0]
    pred = acts[-1][:, 0]
    diff = pred - np.asarray(target, dtype=np.float64)
    loss = float(np.mean(diff * diff))
    delta = (2.0 * diff / n)[:, None]
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        if l < len(layers) - 1:
            delta = delta * (acts[l + 1] > 0)
        gW = acts[l].T @ delta
        gb = delta.sum(axis=0)
        grads.append((gW, gb))
        delta = delta @ W.T
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
