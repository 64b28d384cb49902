"""Seeded expression workloads.

`uneven_dags(T, n, seed)` builds the batch of uneven fused expressions: each expression has
3 to 12 operations drawn from {+, -, *, /, exp, tanh, max}, an output of log-uniform size in
[1e3, 4e6] FP32 elements factored into a rank-1 to rank-4 shape, and 2 to 4 distinct leaves,
each read directly, through a permute view or through a broadcast view (the reference's
leading-dimension broadcast, Tensors.scala:816-855). Leaf data are U(-1, 1) from the Wang
hash (one seed per leaf). Everything is determined by `seed`.
"""
import math

import numpy as np


class Expr(object):
    """One expression of a batch: its tensor, leaf data and byte count."""

    def __init__(self, index, tensor, leaves, leaf_data, ops, view_kinds):
        self.index = index
        self.tensor = tensor
        self.shape = tuple(tensor.shape)
        self.leaves = leaves            # leaf tensors in argument order
        self.leaf_data = leaf_data      # numpy arrays, same order
        self.ops = ops                  # op names in creation order
        self.view_kinds = view_kinds    # "direct" / "permute" / "broadcast" per leaf
        self.out_elements = int(np.prod(self.shape)) if self.shape else 1
        self.leaf_elements = int(sum(d.size for d in leaf_data))

    @property
    def bytes(self):
        """Algorithmic traffic: every distinct leaf read once, the output written once."""
        return 4 * (self.leaf_elements + self.out_elements)


def _factor(n, rank, rng):
    """A shape of `rank` dimensions whose product is close to n (innermost often even)."""
    if rank == 1:
        return (max(1, n),)
    w = rng.dirichlet(np.ones(rank))
    logs = w * math.log(max(n, 1))
    dims = [max(1, int(round(math.exp(x)))) for x in logs]
    if rng.random() < 0.7 and dims[-1] > 1:
        dims[-1] = max(2, dims[-1] // 2 * 2)
    return tuple(dims)


def _uniform(n, seed):
    """U(-1, 1) from the Wang hash, FP64 then rounded (the synthetic-input definition)."""
    i = np.arange(n, dtype=np.uint64).astype(np.uint32) ^ np.uint32(seed & 0xFFFFFFFF)
    with np.errstate(over="ignore"):
        v = (i ^ np.uint32(61)) ^ (i >> np.uint32(16))
        v = v * np.uint32(9)
        v = v ^ (v << np.uint32(4))
        v = v * np.uint32(0x27D4EB2D)
        v = v ^ (v >> np.uint32(15))
    return (-1.0 + 2.0 * (v.astype(np.float64) * (1.0 / 4294967296.0))).astype(np.float32)


OPS = ("+", "-", "*", "/", "exp", "tanh", "max")


def uneven_dags(T, n=512, seed=5001, min_elements=1_000, max_elements=4_000_000):
    rng = np.random.default_rng(seed)
    exprs = []
    for e in range(n):
        size = int(round(math.exp(rng.uniform(math.log(min_elements), math.log(max_elements)))))
        rank = int(rng.integers(1, 5))
        shape = _factor(size, rank, rng)
        nleaves = int(rng.integers(2, 5))
        leaves, data, views, kinds = [], [], [], []
        for k in range(nleaves):
            kind = rng.choice(["direct", "permute", "broadcast"], p=[0.4, 0.3, 0.3])
            if kind == "permute" and rank > 1:
                perm = list(rng.permutation(rank))
                src_shape = tuple(shape[p] for p in perm)
                inv = list(np.argsort(perm))
                arr = _uniform(int(np.prod(src_shape)), seed * 1000 + e * 10 + k).reshape(src_shape)
                t = T.Tensor(arr)
                views.append(t.permute(inv))
            elif kind == "broadcast" and rank > 1:
                keep = int(rng.integers(1, rank))  # leading dimensions kept, the rest appended
                src_shape = tuple(shape[:keep])
                if rng.random() < 0.5:  # a size-1 dimension inside the kept prefix
                    j = int(rng.integers(0, keep))
                    src_shape = tuple(1 if i == j else s for i, s in enumerate(src_shape))
                arr = _uniform(int(np.prod(src_shape)), seed * 1000 + e * 10 + k).reshape(src_shape)
                t = T.Tensor(arr)
                views.append(t.broadcast(shape))
            else:
                kind = "direct"
                arr = _uniform(int(np.prod(shape)), seed * 1000 + e * 10 + k).reshape(shape)
                t = T.Tensor(arr)
                views.append(t)
            leaves.append(t)
            data.append(arr)
            kinds.append(str(kind))
        nops = int(rng.integers(3, 13))
        ops = []
        expr = views[0]
        pool = list(views)
        for _ in range(nops):
            op = OPS[int(rng.integers(0, len(OPS)))]
            ops.append(op)
            if op == "exp":
                expr = T.Tensor.exp(expr)
            elif op == "tanh":
                expr = T.Tensor.tanh(expr)
            else:
                other = pool[int(rng.integers(0, len(pool)))]
                if rng.random() < 0.5:
                    a, b = expr, other
                else:
                    a, b = other, expr
                if op == "max":
                    expr = T.Tensor.max(a, b)
                else:
                    expr = {"+": a.__add__, "-": a.__sub__, "*": a.__mul__, "/": a.__truediv__}[op](b)
            pool.append(expr)
        # arguments in first-visit order of the closure (the kernel's leaf order)
        from . import trees
        order = [p.id for p in trees.parameters(expr.closure)]
        used = [leaves.index(t) for t in order]
        exprs.append(Expr(e, expr, [leaves[i] for i in used], [data[i] for i in used], ops,
                          [kinds[i] for i in used]))
    return exprs


def convolute(T, inp, weight, bias):
    """The convolution of the reference's benchmark (benchmarks.scala:464-548), written with
    the Tensor API as a user would: input [batch, height, width, depth], weight [kh, kw,
    depth, filters], bias [filters]; each output channel is the bias plus the sum, in the
    fold's order, of every translated input channel times its weight broadcast. Returns the
    join of the channels ([batch, height, width, filters])."""
    from functools import reduce
    batch, height, width, depth = inp.shape
    kh, kw, depth2, filters = weight.shape
    if depth2 != depth or tuple(bias.shape) != (filters,):
        raise ValueError("convolute: shapes do not agree")
    inputs = list(inp.split(3))
    weights = [[[list(d.split(0)) for d in kwd.split(0)] for kwd in khkwd.split(0)] for khkwd in weight.split(3)]
    channels = []
    for per_filter, b in zip(weights, bias.split(0)):
        terms = []
        for dy, per_row in zip(range(-(kh // 2), kh - kh // 2), per_filter):
            for dx, per_pixel in zip(range(-(kw // 2), kw - kw // 2), per_row):
                for channel, wt in zip(inputs, per_pixel):
                    terms.append(channel.translate([0, dy, dx]) * wt.broadcast((batch, height, width)))
        channels.append(b.broadcast((batch, height, width)) + reduce(lambda x, y: x + y, terms))
    return T.Tensor.join(channels)


class JoinExpr(object):
    """One join of a split workload: its tensor (a Tensor.join), the leaf tensors and their data
    in argument order, and a description of its shape."""

    def __init__(self, index, tensor, leaves, leaf_data, desc):
        self.index = index
        self.tensor = tensor
        self.leaves = leaves
        self.leaf_data = leaf_data
        self.desc = desc
        self.shape = tuple(tensor.shape)
        self.parts = len(tensor._parts)
        self.out_elements = int(np.prod(self.shape))
        self.leaf_elements = int(sum(d.size for d in leaf_data))


def conv_joins(T, n=64, seed=6001):
    """Joins with many terms per part, the shape of the convolution idiom (convolute): input
    [batch, height, width, depth], a kh x kw kernel with kh, kw in {1, 3, 5}, 2 to 16 filters
    (the parts of the join), so a part holds kh*kw*depth products, at most 320 in a join (its
    one-kernel form compiles in seconds; hiprtc's time grows faster than the body). Sizes are
    drawn from `seed`; data are U(-1, 1) from the Wang hash."""
    rng = np.random.default_rng(seed)
    out = []
    for e in range(n):
        depth = int(rng.choice([1, 2, 4, 8]))
        k = int(rng.choice([1, 3, 5], p=[0.2, 0.6, 0.2]))
        filters = int(rng.choice([2, 4, 8, 16]))
        while k * k * depth * filters > 320:
            if depth > 1:
                depth //= 2
            else:
                filters //= 2
        pixels = int(round(math.exp(rng.uniform(math.log(2e4), math.log(1e6)))))
        hw = max(4, int(math.sqrt(pixels / max(1, int(rng.integers(1, 9))))))
        batch = max(1, pixels // (hw * hw))
        x = _uniform(batch * hw * hw * depth, seed * 1000 + e * 10).reshape(batch, hw, hw, depth)
        w = _uniform(k * k * depth * filters, seed * 1000 + e * 10 + 1).reshape(k, k, depth, filters)
        b = _uniform(filters, seed * 1000 + e * 10 + 2)
        tx, tw, tb = T.Tensor(x), T.Tensor(w), T.Tensor(b)
        joined = convolute(T, tx, tw, tb)
        from . import trees
        clo = trees.Concatenate([p.closure for p in joined._parts])
        order = [p.id for p in trees.parameters(clo)]
        by_id = {id(tx): (tx, x), id(tw): (tw, w), id(tb): (tb, b)}
        leaves = [by_id[id(t)] for t in order]
        out.append(JoinExpr(e, joined, [l[0] for l in leaves], [l[1] for l in leaves],
                            {"batch": batch, "hw": hw, "depth": depth, "kernel": k, "filters": filters}))
    return out
