"""Materialises a facade Tensor graph on the CPU with the reference's programs (test
infrastructure).

Every buffer the reference would create is computed here with refcl-host or the
restated fixed programs, following the same fusion barriers (Tensors.scala:1291-1440):
inline closures run as one refcl kernel over their leaves; leaves are materialised
recursively; `sum` runs the CPU device's sequential program (Tensors.scala:682-696).
Tensor kinds are recognised by class name, so nothing in the product is executed.
"""
import numpy as np

import oracle
from oracle import refcl


def _prod(shape):
    p = 1
    for s in shape:
        p *= int(s)
    return p


def materialize(tensor, mode="precise", memo=None):
    """Flat float32 contents of `tensor` (the buffer a doBuffer would hold)."""
    if memo is None:
        memo = {}
    key = id(tensor)
    if key in memo:
        return memo[key]
    kind = type(tensor).__name__
    n = _prod(tensor.shape)
    if kind == "_DataTensor":
        out = np.asarray(tensor._data, dtype=np.float32).ravel()
    elif kind in ("InlineTensor", "TransformedTensor", "FillTensor"):
        out = run_closure(tensor.closure, tensor.shape, mode, memo)
    elif kind == "_JoinTensor":
        closure = type("Concatenate", (), {})()
        closure.elements = tuple(p.closure for p in tensor._parts)
        out = run_closure(closure, tensor._head_shape, mode, memo)
    elif kind == "_ForwardingTensor":
        out = materialize(tensor._source, mode, memo)
    elif kind == "CachedTensor":
        out = materialize(tensor._source, mode, memo)
    elif kind == "_RandomTensor":
        out = oracle.random_normal(n, tensor.seed) if tensor.normal else oracle.random(n, tensor.seed)
    elif kind == "_ReduceTensor":
        src = materialize(tensor._source, mode, memo)[:_prod(tensor._source.shape)]
        if tensor._kind == "sum":
            out = np.array([oracle.sequential_sum(src)], dtype=np.float32)
        else:
            out = np.array([np.fmax.reduce(src) if src.size else np.nan], dtype=np.float32)
    elif kind == "_RowTensor":
        rows, cols = tensor._source.shape
        src = materialize(tensor._source, mode, memo)[:rows * cols].reshape(rows, cols)
        if tensor._kind == "col":
            out = src.T.copy()
            s, _ = oracle.row_fold(out)
            out = s
        else:
            s, m = oracle.row_fold(src)
            out = s if tensor._kind == "sum" else m
    elif kind == "_GemmTensor":
        a = materialize(tensor._a, mode, memo).reshape(tensor._a.shape)
        b = materialize(tensor._b, mode, memo).reshape(tensor._b.shape)
        out = oracle.gemm_seqk(a, b).ravel()
    else:
        raise TypeError("no CPU materialisation for " + kind)
    memo[key] = out
    return out


def run_closure(closure, shape, mode="precise", memo=None):
    k = refcl.Kernel(closure, shape, mode)
    leaves = [materialize(p.id, mode, memo) for p in k.leaves]
    return k(leaves)


def to_string(tensor, mode="precise"):
    data = materialize(tensor, mode)[:_prod(tensor.shape)]
    return refcl_string(data, tensor.shape)


def refcl_string(flat, shape):
    items = [refcl.java_float_to_string(v) for v in np.asarray(flat, np.float32).ravel()]
    shape = list(shape)
    if not shape:
        return items[0]

    def rec(dims, xs):
        if len(dims) == 1:
            return "[" + ",".join(xs) + "]"
        g = len(xs) // dims[0] if dims[0] else 0
        return "[" + ",".join(rec(dims[1:], xs[i * g:(i + 1) * g]) for i in range(dims[0])) + "]"

    return rec(shape, items)
