"""Learned cost model for fused-expression kernels.

A small MLP (features -> 64 -> 64 -> 64 -> 1, FP32, ReLU) predicts log(microseconds) of one
launch of a fused kernel from features of its expression DAG and its launch configuration
(vector width V, vectors per thread U, workgroup size). It is trained on the device by the
library's own forward, backward and Adam kernels (cs_mlp_*) on timings that the exhaustive
autotuner collects, and it picks each expression's configuration; exhaustive search over the
same candidates remains the fallback and the measure of the model's regret.
"""
import ctypes
import math
import os

import numpy as np

from . import batch as batchmod
from ._capi import check, lib

VECTOR_WIDTHS = (1, 2, 4)
ITEMS = (1, 2, 4)
BLOCKS = (64, 128, 256, 512, 1024)
OPS = ("+", "-", "*", "/", "exp", "tanh", "max")


def configs():
    return [{"vector_width": v, "items_per_thread": u, "block_size": b}
            for v in VECTOR_WIDTHS for u in ITEMS for b in BLOCKS]


def features(expr, cfg):
    """Numeric description of (expression, launch configuration)."""
    shape = expr.shape or (1,)
    v = batchmod.effective_vector_width(expr.shape, cfg["vector_width"])
    u, b = cfg["items_per_thread"], cfg["block_size"]
    n = expr.out_elements
    vectors = max(n // v, 1)
    blocks = max(1.0, vectors / float(b * u))
    f = [
        math.log2(max(n, 1)),
        float(len(shape)),
        math.log2(max(shape[-1], 1)),
        1.0 if shape[-1] % 4 == 0 else 0.0,
        float(len(expr.leaves)),
        float(sum(k == "direct" for k in expr.view_kinds)),
        float(sum(k == "permute" for k in expr.view_kinds)),
        float(sum(k == "broadcast" for k in expr.view_kinds)),
        math.log2(max(expr.leaf_elements, 1)),
        float(len(expr.ops)),
    ]
    f += [float(expr.ops.count(o)) for o in OPS]
    f += [float(v), float(u), math.log2(b), math.log2(blocks), math.log2(blocks / 2048.0 + 1.0),
          1.0 if v != cfg["vector_width"] else 0.0]
    return np.array(f, dtype=np.float32)


class Network(object):
    """The device MLP (cs_mlp)."""

    def __init__(self, tensors, widths, seed=1):
        self.T = tensors
        w = (ctypes.c_int32 * len(widths))(*widths)
        h = ctypes.c_void_p()
        check(lib.cs_mlp_create(tensors.device, w, len(widths), seed, ctypes.byref(h)), "cs_mlp_create")
        self.handle = h.value
        self.widths = list(widths)
        c = ctypes.c_int64()
        check(lib.cs_mlp_param_count(self.handle, ctypes.byref(c)))
        self.count = c.value
        self.step = 0
        self._grads = tensors._alloc(self.count * 4)
        self._loss = tensors._alloc(4)

    def get_params(self):
        out = np.empty(self.count, np.float32)
        check(lib.cs_mlp_get_params(self.T.stream, self.handle, out.ctypes.data_as(ctypes.c_void_p), self.count))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib.cs_mlp_set_params(self.T.stream, self.handle, p.ctypes.data_as(ctypes.c_void_p), self.count))

    def forward_device(self, x_ptr, n, y_ptr):
        check(lib.cs_mlp_forward(self.T.stream, self.handle, x_ptr, n, y_ptr))

    def gradients_device(self, x_ptr, t_ptr, n):
        """Runs the backward pass; returns (loss, gradient buffer)."""
        check(lib.cs_mlp_backward(self.T.stream, self.handle, x_ptr, t_ptr, n, self._loss.ptr, self._grads.ptr))
        return self._loss, self._grads

    def adam(self, lr):
        self.step += 1
        check(lib.cs_mlp_adam(self.T.stream, self.handle, self._grads.ptr, lr, self.step))

    def release(self):
        if self.handle:
            lib.cs_mlp_release(self.handle)
            self._grads.release()
            self._loss.release()
            self.handle = None


class CostModel(object):
    """Predicts log-microseconds (relative to the expression's mean when trained with groups);
    selects the fastest predicted configuration."""

    def __init__(self, tensors, nfeatures, hidden=64, layers=3, seed=1):
        self.T = tensors
        self.net = Network(tensors, [nfeatures] + [hidden] * layers + [1], seed)
        self.mean = np.zeros(nfeatures, np.float32)
        self.std = np.ones(nfeatures, np.float32)
        self.ymean = 0.0
        self.history = []

    def _upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        b = self.T._alloc(max(a.nbytes, 4))
        check(lib.cs_buffer_write(self.T.stream, b.handle, 0, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1))
        return b

    def fit(self, X, us, epochs=300, batch=512, lr=2e-3, seed=0, groups=None):
        """Trains on features X[n, f] and measured microseconds; returns the loss history.
        With `groups` (the expression of each row), the target is each time's log relative to
        the mean log-time of its expression: the choice among one expression's configurations
        depends only on those differences, and the spread of sizes across expressions no
        longer dominates the loss."""
        X = np.asarray(X, np.float32)
        y = np.log(np.asarray(us, np.float64))
        if groups is not None:
            groups = np.asarray(groups)
            for g in np.unique(groups):
                sel = groups == g
                y[sel] -= y[sel].mean()
        y = y.astype(np.float32)
        self.mean = X.mean(axis=0)
        self.std = X.std(axis=0) + 1e-3
        self.ymean = float(y.mean())
        Xn = (X - self.mean) / self.std
        yn = y - self.ymean
        rng = np.random.default_rng(seed)
        n = len(Xn)
        for ep in range(epochs):
            perm = rng.permutation(n)
            xb = self._upload(Xn[perm])
            yb = self._upload(yn[perm])
            losses = []
            for s in range(0, n - batch + 1, batch):
                loss_buf, _ = self.net.gradients_device(xb.ptr + s * X.shape[1] * 4, yb.ptr + s * 4, batch)
                self.net.adam(lr * (0.5 if ep > epochs * 0.7 else 1.0))
                if s == 0:
                    losses.append(float(loss_buf.to_numpy(1)[0]))
            xb.release()
            yb.release()
            self.history.append(float(np.mean(losses)) if losses else float("nan"))
        return self.history

    def predict_log_us(self, X):
        X = (np.asarray(X, np.float32) - self.mean) / self.std
        xb = self._upload(X)
        out = self.T._alloc(len(X) * 4)
        self.net.forward_device(xb.ptr, len(X), out.ptr)
        y = out.to_numpy(len(X)) + self.ymean
        xb.release()
        out.release()
        return y

    def choose(self, expr, candidates):
        X = np.stack([features(expr, c) for c in candidates])
        return candidates[int(np.argmin(self.predict_log_us(X)))]

    def release(self):
        self.net.release()


MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "costmodel.npz")


def load(tensors, path=MODEL_PATH):
    """The trained model saved by the exhaustive sweep (parameters and feature scaling), on
    `tensors`'s device; None when no parameters were saved."""
    if not os.path.exists(path):
        return None
    d = np.load(path)
    model = CostModel(tensors, int(d["widths"][0]), hidden=int(d["widths"][1]), layers=len(d["widths"]) - 2)
    model.net.set_params(d["params"])
    model.mean, model.std, model.ymean = d["mean"], d["std"], float(d["ymean"][0])
    return model


def measure(eset, fused_by_variant, cfgs, reps=3, min_us=40.0, progress=None):
    """Exhaustive autotuner sweep: microseconds of one launch of expression i with each
    configuration, from a graph of back-to-back launches. Returns {(i, cfg_index): us}."""
    out = {}
    for i, e in enumerate(eset.exprs):
        if progress:
            progress(i)
        for ci, cfg in enumerate(cfgs):
            f = fused_by_variant[(cfg["vector_width"], cfg["items_per_thread"])][i]
            est = max(2.0, e.bytes / 4e6)  # ~4 TB/s guess to size the repetition count
            k = int(min(64, max(4, min_us / est)))
            g = batchmod.Graph(eset.T, lambda: [eset.launch(i, f, cfg["block_size"]) for _ in range(k)])
            out[(i, ci)] = g.time_us(reps) / k
            g.release()
    return out


# ------------------------------------------------------------------------ fusion splits
def _counts(closure):
    """(nodes, products, sums, leaves, views) of a closure, each shared node counted once."""
    from . import trees
    nodes, _, leaves = trees.serialize(closure)
    ops = [n.op for n in nodes]
    from ._capi import OP_MUL, OP_ADD, OP_TRANSFORM
    return (len(nodes), ops.count(OP_MUL), ops.count(OP_ADD), len(leaves), ops.count(OP_TRANSFORM))


def split_features(join, groups):
    """Numeric description of (join, number of kernels it is cut into)."""
    from . import trees
    parts = len(join._parts)
    head = tuple(join._head_shape)
    n = 1
    for s in head:
        n *= s
    whole = _counts(trees.Concatenate([p.closure for p in join._parts]))
    heads, per = join.closures(groups)
    one = _counts(heads[0])
    f = [
        math.log2(max(n, 1)), float(len(head)), math.log2(max(head[-1], 1)),
        float(parts), math.log2(parts), float(groups), math.log2(groups), float(per),
        math.log2(whole[0]), math.log2(max(whole[1], 1)), float(whole[3]), float(whole[4]),
        math.log2(one[0]), math.log2(max(one[1], 1)), math.log2(max(one[1], 1) / per),
        math.log2(max(n, 1) * parts * 4.0), math.log2(max(n, 1)) + math.log2(max(one[1], 1)) + math.log2(groups),
    ]
    return np.array(f, dtype=np.float32)


class SplitModel(CostModel):
    """The cost model of fusion splits: predicts the log-time of a join cut into g kernels
    (relative to the join's mean, as CostModel with groups) and picks g. A cut is taken only
    when it is predicted faster than one kernel by more than `margin` (a fraction, calibrated
    on the training joins): one kernel is the reference's method and usually the fastest."""

    margin = 0.0

    def predicted(self, join, largest=8):
        from . import autotune
        cands = autotune.split_candidates(len(join._parts), largest)
        X = np.stack([split_features(join, g) for g in cands])
        return cands, self.predict_log_us(X)

    def choose_groups(self, join, largest=8):
        cands, pred = self.predicted(join, largest)
        return pick_groups(cands, pred, self.margin)


def pick_groups(cands, pred, margin):
    """The cut with the lowest predicted log-time, if it beats one kernel (cands[0] == 1) by
    more than `margin`; otherwise one kernel."""
    best = int(np.argmin(pred))
    if cands[best] != 1 and pred[best] < pred[0] - math.log1p(margin):
        return cands[best]
    return 1


SPLIT_MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "splitmodel.npz")


def save(model, path):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=model.net.get_params(), mean=model.mean, std=model.std,
             ymean=np.array([model.ymean]), widths=np.array(model.net.widths),
             margin=np.array([getattr(model, "margin", 0.0)]))


def load_split_model(tensors, path=SPLIT_MODEL_PATH):
    """The shipped split model on `tensors`'s device; None when none was saved."""
    if not os.path.exists(path):
        return None
    d = np.load(path)
    model = SplitModel(tensors, int(d["widths"][0]), hidden=int(d["widths"][1]), layers=len(d["widths"]) - 2)
    model.net.set_params(d["params"])
    model.mean, model.std, model.ymean = d["mean"], d["std"], float(d["ymean"][0])
    if "margin" in d:
        model.margin = float(d["margin"][0])
    return model
