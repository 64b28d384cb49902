"""The Tensor API of Compute.scala over the gfx950 back-end.

Mirror of `trait Tensors` (Tensors.scala:199-1442): lazy, immutable N-dimensional FP32
arrays whose elementwise expressions and affine views fuse into one kernel per
materialisation. The classes and their names follow the reference:

  * InlineTensor       -- a closure that is inlined into the consumer's kernel
  * TransformedTensor  -- an affine view (checkpoint + matrix) folded into index math
  * FillTensor         -- a constant
  * NonInlineTensor    -- a buffer-backed leaf (data, random, join, sum, reshape, nonInline)
  * CachedTensor       -- a leaf holding a device buffer until `release` (doCache)

Slow actions (toString, flatArray, read*, doCache, doBuffer) compile the closure through
the C ABI (structural kernel cache in the library), materialise the leaves, launch on
the object's stream and read back through pinned staging.

Differences from the reference, on purpose:
  * `%` computes fmodf (the reference emits float `%`, which OpenCL C rejects);
  * +-infinity literals compile (the reference emits the undefined `INFINITE`);
  * `sum` is a correct tree for every length (the reference's parallel program drops
    elements for non power-of-two local sizes, Tensors.scala:375-387);
  * the kernel cache key includes the output shape (the reference keys on the closure
    alone, so two fills of one value and different ranks share a kernel).
"""
import ctypes
import struct
import threading
from collections.abc import Sequence

import numpy as np

from . import _capi as capi
from . import affine
from . import jvm_format
from . import trees
from ._capi import check, lib


class IllegalArgumentException(ValueError):
    """The JVM exception the reference raises for shape errors."""


def _product(shape):
    p = 1
    for s in shape:
        p *= int(s)
    return p


def _f32(v):
    return struct.unpack("<f", struct.pack("<f", float(v)))[0]


# ============================================================================ buffers
class PendingBuffer(object):
    """A device buffer reference (PendingBuffer / DeviceBuffer, Tensors.scala:253-299,
    OpenCL.scala:614-717). Work producing it is queued on the owner's stream."""

    __slots__ = ("tensors", "handle", "ptr", "nbytes", "_released")

    def __init__(self, tensors, handle, nbytes):
        self.tensors = tensors
        self.handle = handle
        p = ctypes.c_void_p()
        check(lib.cs_buffer_device_ptr(handle, ctypes.byref(p)), "cs_buffer_device_ptr")
        self.ptr = p.value
        self.nbytes = nbytes
        self._released = False

    def retain(self):
        check(lib.cs_buffer_retain(self.handle), "cs_buffer_retain")
        return PendingBuffer(self.tensors, self.handle, self.nbytes)

    def release(self):
        if self._released:
            raise capi.InvalidMemObject("buffer reference released twice")
        self._released = True
        check(lib.cs_buffer_release(self.handle), "cs_buffer_release")

    def to_numpy(self, count):
        out = np.empty(count, dtype=np.float32)
        if count:
            check(lib.cs_buffer_read(self.tensors._stream, self.handle, 0,
                                     out.ctypes.data_as(ctypes.c_void_p), count * 4), "cs_buffer_read")
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if not self._released:
            self.release()


# ============================================================================ tensors
class Tensor(object):
    """Base of every tensor (Tensors.scala:636-1261)."""

    _tensors = None
    shape = ()
    padding = 0.0

    # ---------------------------------------------------------------- closure
    @property
    def closure(self):
        raise NotImplementedError

    def getClosure(self):
        return self.closure

    @property
    def arrayTerm(self):
        """array.parameter(this, padding, shape), created once per tensor
        (Tensors.scala:1253-1260)."""
        at = self.__dict__.get("_array_term")
        if at is None:
            at = trees.ArrayParameter(self, self.padding, self.shape)
            self.__dict__["_array_term"] = at
        return at

    # ---------------------------------------------------------------- materialisation
    def doBuffer(self):
        """Materialises this tensor; the caller owns the returned reference."""
        memo = {}
        try:
            return self._buffer(memo)
        finally:
            for b in memo.values():
                b.release()

    def _buffer(self, memo):
        b = memo.get(id(self))
        if b is None:
            b = self._materialize(memo)
            memo[id(self)] = b
        return b.retain()

    def _materialize(self, memo):
        raise NotImplementedError

    # ---------------------------------------------------------------- slow actions
    def flatArray(self):
        n = _product(self.shape)
        with self.doBuffer() as b:
            return b.to_numpy(n)

    flatBuffer = flatArray

    def toString(self):
        return jvm_format.nested_to_string(self.flatArray(), self.shape)

    def __str__(self):
        return self.toString()

    def __repr__(self):
        return "%s(shape=%r)" % (type(self).__name__, tuple(self.shape))

    def readScalar(self):
        return float(self.flatArray()[0])

    def read1DArray(self):
        return [float(v) for v in self.flatArray()]

    def _readND(self, rank):
        a = self.flatArray().reshape(self.shape)
        if len(self.shape) != rank:
            raise IllegalArgumentException("tensor of rank %d read as %dD" % (len(self.shape), rank))
        return a.tolist()

    def read2DArray(self):
        return self._readND(2)

    def read3DArray(self):
        return self._readND(3)

    def read4DArray(self):
        return self._readND(4)

    def read5DArray(self):
        return self._readND(5)

    read1DSeq = read1DArray
    read2DSeq = read2DArray
    read3DSeq = read3DArray
    read4DSeq = read4DArray
    read5DSeq = read5DArray

    def doCache(self):
        """Eagerly materialises into a CachedTensor that keeps its buffer until
        `release()` (Tensors.scala:642-666)."""
        buf = self.doBuffer()
        return CachedTensor(self._tensors, self.shape, self.padding, buf, source=self)

    # ---------------------------------------------------------------- delayed operators
    @property
    def nonInline(self):
        raise NotImplementedError

    def _derived(self, closure):
        return InlineTensor(self._tensors, self.shape, self.padding, closure)

    def _binary(self, op, rhs):
        if not isinstance(rhs, Tensor):
            return NotImplemented
        left, right = self._tensors._auto_broadcast(self, rhs)
        return left._derived(trees.Binary(op, left.closure, right.closure))

    def __add__(self, rhs):
        return self._binary("+", rhs)

    def __sub__(self, rhs):
        return self._binary("-", rhs)

    def __mul__(self, rhs):
        return self._binary("*", rhs)

    def __truediv__(self, rhs):
        return self._binary("/", rhs)

    def __mod__(self, rhs):
        return self._binary("%", rhs)

    def __neg__(self):
        return self._derived(trees.Unary("neg", self.closure))

    def __pos__(self):
        return self

    def broadcast(self, new_shape):
        """(Tensors.scala:816-855) Old dimension i maps to new dimension i; new trailing
        dimensions are appended, and size-1 dimensions stretch."""
        new_shape = tuple(int(s) for s in new_shape)
        if new_shape == tuple(self.shape):
            return self
        if isinstance(self, FillTensor):
            return FillTensor(self._tensors, new_shape, self.padding, self.value)
        return self._broadcast_transform(new_shape)

    def _broadcast_transform(self, new_shape):
        n_new, n = len(new_shape), len(self.shape)
        if n > n_new:
            raise IllegalArgumentException("Cannot broadcast %s to %s" % (list(self.shape), list(new_shape)))
        m = [0.0] * ((n_new + 1) * n)
        for i in range(n):
            if self.shape[i] == new_shape[i]:
                m[i * (n_new + 1) + i] = 1.0
            elif self.shape[i] == 1:
                pass
            else:
                raise IllegalArgumentException(
                    "Cannot broadcast [%s] to [%s]" % (",".join(map(str, self.shape)), ",".join(map(str, new_shape))))
        return self._transform(new_shape, m)

    def reshape(self, new_shape):
        """A fusion barrier sharing this tensor's buffer (Tensors.scala:879-888)."""
        new_shape = tuple(int(s) for s in new_shape)
        if _product(new_shape) != _product(self.shape):
            raise IllegalArgumentException("reshape from %s to %s" % (list(self.shape), list(new_shape)))
        return _ForwardingTensor(self._tensors, new_shape, self.padding, self)

    def scale(self, new_shape):
        """(Tensors.scala:950-965)"""
        new_shape = tuple(int(s) for s in new_shape)
        n = len(new_shape)
        if n != len(self.shape):
            raise IllegalArgumentException("scale must keep the rank")
        m = [0.0] * (n * (n + 1))
        for i in range(n):
            m[i * (n + 1) + i] = float(self.shape[i]) / new_shape[i]
        return self._transform(new_shape, m)

    def translate(self, offset, new_shape=None):
        """out[i] = this[i - offset], padding outside (Tensors.scala:970-976)."""
        if len(offset) != len(self.shape):
            raise IllegalArgumentException("offset rank differs from the tensor rank")
        shape = tuple(self.shape) if new_shape is None else tuple(int(s) for s in new_shape)
        return self._transform(shape, affine.translate([-float(o) for o in offset]))

    def _transform(self, new_shape, matrix1):
        """(Tensors.scala:978-1003) Views of a view compose into one matrix per checkpoint."""
        if isinstance(self, TransformedTensor):
            m = affine.pre_concatenate(matrix1, self.matrix, len(new_shape))
            return TransformedTensor(self._tensors, new_shape, self.padding, self.checkpoint, m)
        return TransformedTensor(self._tensors, new_shape, self.padding, self, matrix1)

    def permute(self, dimensions):
        """(Tensors.scala:1008-1025) new dimension i is old dimension dimensions[i]."""
        n = len(self.shape)
        if len(dimensions) != n:
            raise IllegalArgumentException("permute needs %d dimensions" % n)
        new_shape = [0] * n
        m = [0.0] * (n * (n + 1))
        for new_i in range(n):
            old_i = dimensions[new_i]
            new_shape[new_i] = self.shape[old_i]
            m[old_i * (n + 1) + new_i] = 1.0
        return self._transform(tuple(new_shape), m)

    def transpose(self):
        return self.permute(list(reversed(range(len(self.shape)))))

    def split(self, dimension):
        """Views along `dimension`, without copies (Tensors.scala:1035-1074)."""
        return TensorSeq(self, dimension)

    def sum(self):
        """Full FP32 sum (Tensors.scala:673-771)."""
        return _ReduceTensor(self._tensors, self, "sum")

    # Additions without a reference API (the reference writes them as split/reduce chains).
    def maxValue(self):
        return _ReduceTensor(self._tensors, self, "max")

    def rowSum(self):
        return _RowTensor(self._tensors, self, "sum")

    def rowMax(self):
        return _RowTensor(self._tensors, self, "max")

    def colSum(self):
        return _RowTensor(self._tensors, self, "col")


class InlineTensor(Tensor):
    """(Tensors.scala:1400-1411)"""

    def __init__(self, tensors, shape, padding, closure):
        self._tensors = tensors
        self.shape = tuple(shape)
        self.padding = _f32(padding)
        self._closure = closure

    @property
    def closure(self):
        return self._closure

    def _materialize(self, memo):
        return self._tensors._enqueue_closure(self.closure, self.shape, memo)

    @property
    def nonInline(self):
        return _ForwardingTensor(self._tensors, self.shape, self.padding, self)


class TransformedTensor(InlineTensor):
    """(Tensors.scala:1413-1428) closure = checkpoint.arrayTerm.transform(matrix).extract"""

    def __init__(self, tensors, shape, padding, checkpoint, matrix):
        self._tensors = tensors
        self.shape = tuple(shape)
        self.checkpoint = checkpoint
        self.matrix = list(matrix)
        self.padding = _f32(checkpoint.padding if padding is None else padding)
        self._closure = None

    @property
    def closure(self):
        if self._closure is None:
            self._closure = trees.Extract(trees.Transform(self.checkpoint.arrayTerm, self.matrix, len(self.shape)))
        return self._closure


class FillTensor(InlineTensor):
    """(Tensors.scala:1394-1397)"""

    def __init__(self, tensors, shape, padding, value):
        InlineTensor.__init__(self, tensors, shape, padding, trees.FloatLiteral(value))
        self.value = _f32(value)


class NonInlineTensor(Tensor):
    """(Tensors.scala:1431-1440) closure = arrayTerm.extract"""

    @property
    def closure(self):
        c = self.__dict__.get("_closure")
        if c is None:
            c = trees.Extract(self.arrayTerm)
            self.__dict__["_closure"] = c
        return c

    @property
    def nonInline(self):
        return self


class CachedTensor(NonInlineTensor):
    """A tensor whose buffer is kept until `release()` (Tensors.scala:602, 642-666)."""

    def __init__(self, tensors, shape, padding, buffer, source=None):
        self._tensors = tensors
        self.shape = tuple(shape)
        self.padding = _f32(padding)
        self._buf = buffer
        self._source = source  # the tensor this cache was taken from

    def _materialize(self, memo):
        if self._buf is None:
            raise capi.InvalidMemObject("the cache of this tensor has been released")
        return self._buf.retain()

    def release(self):
        if self._buf is not None:
            b, self._buf = self._buf, None
            b.release()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()

    @property
    def buffer(self):
        return self._buf


class _DataTensor(NonInlineTensor):
    """Tensor(elements) (Tensors.scala:445-463): uploaded through pinned staging."""

    def __init__(self, tensors, data, shape, padding):
        self._tensors = tensors
        self.shape = tuple(shape)
        self.padding = _f32(padding)
        self._data = np.ascontiguousarray(data, dtype=np.float32).ravel()

    def _materialize(self, memo):
        t = self._tensors
        buf = t._alloc(self._data.size * 4)
        if self._data.size:
            check(lib.cs_buffer_write(t._stream, buf.handle, 0, self._data.ctypes.data_as(ctypes.c_void_p),
                                      self._data.size * 4, 0), "cs_buffer_write")
        return buf


class _ForwardingTensor(NonInlineTensor):
    """nonInline and reshape: a barrier that shares the source buffer (Tensors.scala:879-888,
    1405-1410)."""

    def __init__(self, tensors, shape, padding, source):
        self._tensors = tensors
        self.shape = tuple(shape)
        self.padding = _f32(padding)
        self._source = source

    def _materialize(self, memo):
        return self._source._buffer(memo)


class _RandomTensor(NonInlineTensor):
    """Tensor.random / Tensor.randomNormal (Tensors.scala:479-524)."""

    def __init__(self, tensors, shape, padding, seed, normal):
        self._tensors = tensors
        self.shape = tuple(shape)
        self.padding = _f32(padding)
        self.seed = int(seed) & 0xFFFFFFFF
        self.normal = normal

    def _materialize(self, memo):
        t = self._tensors
        n = _product(self.shape)
        if self.normal:
            padded = n + (n & 1)
            buf = t._alloc(padded * 4)
            check(lib.cs_random_normal(t._stream, buf.ptr, n, self.seed), "cs_random_normal")
        else:
            buf = t._alloc(n * 4)
            check(lib.cs_random(t._stream, buf.ptr, n, self.seed), "cs_random")
        return buf


class _JoinTensor(NonInlineTensor):
    """Tensor.join(seq): one kernel writing n floats per coordinate (Tensors.scala:577-598)."""

    def __init__(self, tensors, parts):
        head = parts[0]
        self._tensors = tensors
        self.shape = tuple(head.shape) + (len(parts),)
        self.padding = head.padding
        self._parts = list(parts)
        self._head_shape = tuple(head.shape)

    def closures(self, groups=1):
        """The kernels that materialise this join when it is cut into `groups` kernels of
        consecutive parts (groups divides the number of parts): one Concatenate per group,
        then, for groups > 1, the kernel that gathers the groups' outputs into the joined
        layout (its leaves are the group outputs, given as (shape, padding) here)."""
        per = len(self._parts) // groups
        heads = [trees.Concatenate([p.closure for p in self._parts[k * per:(k + 1) * per]]) for k in range(groups)]
        return heads, per

    def _materialize(self, memo):
        t = self._tensors
        groups = t._join_groups(self)
        if groups == 1:
            closure = trees.Concatenate([p.closure for p in self._parts])
            return t._enqueue_closure(closure, self._head_shape, memo)
        # A fusion split: each group of parts is its own kernel writing [head, per]; a last
        # kernel gathers them into [head, parts]. Only copies are added, so the bits are the
        # one-kernel join's.
        heads, per = self.closures(groups)
        parts = []
        try:
            for h in heads:
                parts.append(CachedTensor(t, self._head_shape + (per,), self.padding,
                                          t._enqueue_closure(h, self._head_shape, memo)))
            r = len(self._head_shape)
            gather = [trees.Extract(trees.Transform(p.arrayTerm, _split_matrix(r + 1, r, j), r))
                      for p in parts for j in range(per)]
            return t._enqueue_closure(trees.Concatenate(gather), self._head_shape, memo)
        finally:
            for p in parts:
                p.release()


class _ReduceTensor(NonInlineTensor):
    def __init__(self, tensors, source, kind):
        self._tensors = tensors
        self.shape = ()
        self.padding = source.padding
        self._source = source
        self._kind = kind

    def _fusable(self, memo):
        t = self._tensors
        s = self._source
        return (self._kind == "sum" and t.fused_sums and isinstance(s, InlineTensor) and id(s) not in memo and
                not isinstance(s.closure, trees.Concatenate))

    def _materialize(self, memo):
        t = self._tensors
        if self._fusable(memo):
            f, leaves = t.compile_sum(self._source.closure, self._source.shape)
            try:
                args = [leaf.id._buffer(memo) for leaf in leaves]
                try:
                    out = t._alloc(4)
                    ptrs = (ctypes.c_void_p * max(len(args), 1))(*[a.ptr for a in args])
                    check(lib.cs_fused_sum(t._stream, f, ptrs, len(args), out.ptr, None), "cs_fused_sum")
                    t.launches += 1
                    return out
                finally:
                    for a in args:
                        a.release()
            finally:
                lib.cs_fused_release(f)
        src = self._source._buffer(memo)
        try:
            out = t._alloc(4)
            n = _product(self._source.shape)
            if self._kind == "sum":
                check(lib.cs_reduce_sum(t._stream, src.ptr, n, out.ptr, None), "cs_reduce_sum")
            else:
                check(lib.cs_reduce_max(t._stream, src.ptr, n, out.ptr), "cs_reduce_max")
            return out
        finally:
            src.release()


class _RowTensor(NonInlineTensor):
    def __init__(self, tensors, source, kind):
        if len(source.shape) != 2:
            raise IllegalArgumentException("row and column reductions take a matrix")
        self._tensors = tensors
        self.shape = (source.shape[1],) if kind == "col" else (source.shape[0],)
        self.padding = source.padding
        self._source = source
        self._kind = kind

    def _materialize(self, memo):
        t = self._tensors
        rows, cols = self._source.shape
        src = self._source._buffer(memo)
        try:
            out = t._alloc(max(self.shape[0], 1) * 4)
            if self._kind == "sum":
                check(lib.cs_reduce_rows(t._stream, src.ptr, rows, cols, out.ptr, None, None, None), "cs_reduce_rows")
            elif self._kind == "max":
                check(lib.cs_reduce_rows(t._stream, src.ptr, rows, cols, None, out.ptr, None, None), "cs_reduce_rows")
            else:
                check(lib.cs_reduce_cols(t._stream, src.ptr, rows, cols, out.ptr), "cs_reduce_cols")
            return out
        finally:
            src.release()


class _GemmTensor(NonInlineTensor):
    def __init__(self, tensors, a, b, variant):
        if len(a.shape) != 2 or len(b.shape) != 2 or a.shape[1] != b.shape[0]:
            raise IllegalArgumentException("matrixMultiply of %s and %s" % (list(a.shape), list(b.shape)))
        self._tensors = tensors
        self.shape = (a.shape[0], b.shape[1])
        self.padding = a.padding
        self._a, self._b, self._variant = a, b, variant

    def _materialize(self, memo):
        t = self._tensors
        M, K = self._a.shape
        N = self._b.shape[1]
        a = self._a._buffer(memo)
        b = self._b._buffer(memo)
        try:
            out = t._alloc(max(M * N, 1) * 4)
            check(lib.cs_gemm(t._stream, self._variant, M, N, K, a.ptr, K, b.ptr, N, out.ptr, N), "cs_gemm")
            return out
        finally:
            a.release()
            b.release()


def _split_matrix(n, d, index):
    """The view of slice `index` along dimension d of a rank-n tensor: n source rows x
    (n - 1 + 1) columns (Tensors.scala:1038-1073)."""
    m = [0.0] * (n * n)
    for i in range(d):
        m[i * n + i] = 1.0
    m[d * n + n - 1] = float(index)
    for i in range(d + 1, n):
        m[i * n + i - 1] = 1.0
    return tuple(m)


def _split_fold(closure, shape):
    """Recognises `t.split(d).reduce(_ + _)` and `t.split(d).reduce(Tensor.max)` over a
    materialised tensor t, the way the reference writes axis sums and maxima (README.md:342,
    TensorsSpec.scala:477; a left fold of elementwise ops over the slices, in order).
    Returns (parameter, d, op) or None."""
    if not isinstance(closure, trees.Binary) or closure.op not in ("+", "max"):
        return None
    op = closure.op
    slices = []
    node = closure
    while isinstance(node, trees.Binary) and node.op == op:
        slices.append(node.rhs)
        node = node.lhs
    slices.append(node)
    slices.reverse()
    param = None
    for s in slices:
        if not (isinstance(s, trees.Extract) and isinstance(s.array, trees.Transform) and
                isinstance(s.array.array, trees.ArrayParameter)):
            return None
        if param is None:
            param = s.array.array
        elif s.array.array is not param:
            return None
    n = len(param.shape)
    if n < 1 or len(shape) != n - 1:
        return None
    for d in range(n):
        if param.shape[d] != len(slices) or tuple(param.shape[:d] + param.shape[d + 1:]) != tuple(shape):
            continue
        if all(s.array.matrix == _split_matrix(n, d, i) for i, s in enumerate(slices)):
            return param, d, op
    return None


def _is_view(node, rank, rows):
    """Extract(Transform(ArrayParameter, matrix)) of the given output rank: returns
    (parameter, matrix as rows x (rank + 1) tuples) or None."""
    if not (isinstance(node, trees.Extract) and isinstance(node.array, trees.Transform) and
            isinstance(node.array.array, trees.ArrayParameter) and node.array.rank == rank):
        return None
    m = node.array.matrix
    if len(m) != rows * (rank + 1):
        return None
    return node.array.array, tuple(tuple(m[r * (rank + 1):(r + 1) * (rank + 1)]) for r in range(rows))


def _left_fold_terms(node, op):
    """The operands of a left fold ((t0 op t1) op t2) ..., in order."""
    terms = []
    while isinstance(node, trees.Binary) and node.op == op:
        terms.append(node.rhs)
        node = node.lhs
    terms.append(node)
    terms.reverse()
    return terms


def _match_mm2(product):
    """matrixMultiply2 of the reference (README.md:337-346, TensorsSpec.scala:472-479): the
    tensor split along dimension 1 and folded is A.broadcast([i,j,k]) * B'.broadcast([i,j,k])
    with A an [i,j] matrix and B' = B.reshape([1,j,k]). Returns (A tensor, B' tensor, i, j, k)
    or None. Either operand order is accepted (an FP32 product commutes exactly)."""
    if not isinstance(product, InlineTensor) or len(product.shape) != 3:
        return None
    c = product.closure
    if not (isinstance(c, trees.Binary) and c.op == "*"):
        return None
    i, j, k = product.shape
    for x, y in ((c.lhs, c.rhs), (c.rhs, c.lhs)):
        va, vb = _is_view(x, 3, 2), _is_view(y, 3, 3)
        if va is None or vb is None:
            continue
        (pa, ma), (pb, mb) = va, vb
        if (tuple(pa.shape) == (i, j) and ma == ((1, 0, 0, 0), (0, 1, 0, 0)) and
                tuple(pb.shape) == (1, j, k) and mb == ((0, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0))):
            return pa.id, pb.id, i, j, k
    return None


def _match_mm1(closure, shape):
    """matrixMultiply1 of the reference (README.md:312-325, TensorsSpec.scala:506-518):
    join over the K columns of B of the left fold over j of A.split(1)(j) *
    B.split(1)(k).split(0)(j).broadcast([i]). Returns (A tensor, B tensor, i, j, k) or None."""
    if not isinstance(closure, trees.Concatenate) or len(shape) != 1:
        return None
    i = shape[0]
    pa = pb = None
    J = None
    for kk, part in enumerate(closure.elements):
        terms = _left_fold_terms(part, "+")
        if J is None:
            J = len(terms)
        if len(terms) != J:
            return None
        for jj, t in enumerate(terms):
            if not (isinstance(t, trees.Binary) and t.op == "*"):
                return None
            l, r = _is_view(t.lhs, 1, 2), _is_view(t.rhs, 1, 2)
            if l is None or r is None:
                return None
            if l[1] != ((1, 0), (0, jj)) or r[1] != ((0, jj), (0, kk)):
                return None
            if pa is None:
                pa, pb = l[0], r[0]
            if l[0] is not pa or r[0] is not pb:
                return None
    K = len(closure.elements)
    if tuple(pa.shape) != (i, J) or tuple(pb.shape) != (J, K):
        return None
    return pa.id, pb.id, i, J, K


class TensorSeq(Sequence):
    """The sequence returned by split (Tensors.scala:1038-1073)."""

    def __init__(self, tensor, dimension):
        self._t = tensor
        self._d = int(dimension)
        if not 0 <= self._d < len(tensor.shape):
            raise IllegalArgumentException("split dimension out of range")
        self._new_shape = tuple(s for i, s in enumerate(tensor.shape) if i != self._d)

    def __len__(self):
        return self._t.shape[self._d]

    def __getitem__(self, index):
        if isinstance(index, slice):
            return [self[i] for i in range(*index.indices(len(self)))]
        if index < 0:
            index += len(self)
        if not 0 <= index < len(self):
            raise IndexError(index)
        return self._t._transform(self._new_shape, _split_matrix(len(self._t.shape), self._d, index))

    def __str__(self):
        return "TensorSeq(" + ", ".join(t.toString() for t in self) + ")"

    toString = __str__


# ============================================================================ the back-end
class TensorCompanion(object):
    """`object Tensor` of the reference (Tensors.scala:395-600), bound to one back-end."""

    def __init__(self, tensors):
        self._t = tensors

    def __call__(self, elements, padding=0.0):
        data, shape = _flatten(elements)
        return _DataTensor(self._t, data, shape, padding)

    apply = __call__

    def scalar(self, value, padding=0.0):
        return self.fill(value, (), padding)

    def fill(self, value, shape, padding=0.0):
        return FillTensor(self._t, tuple(int(s) for s in shape), padding, value)

    def random(self, shape, seed=None, padding=0.0):
        if seed is None:
            seed = np.random.randint(0, 2 ** 31)
        return _RandomTensor(self._t, tuple(shape), padding, seed, normal=False)

    def randomNormal(self, shape, seed=None, padding=0.0):
        if seed is None:
            seed = np.random.randint(0, 2 ** 31)
        return _RandomTensor(self._t, tuple(shape), padding, seed, normal=True)

    def _unary(self, op, x):
        return x._derived(trees.Unary(op, x.closure))

    def abs(self, x):
        return self._unary("abs", x)

    def sqrt(self, x):
        return self._unary("sqrt", x)

    def tanh(self, x):
        return self._unary("tanh", x)

    def exp(self, x):
        return self._unary("exp", x)

    def log(self, x):
        return self._unary("log", x)

    def min(self, a, b):
        left, right = self._t._auto_broadcast(a, b)
        return left._derived(trees.Binary("min", left.closure, right.closure))

    def max(self, a, b):
        left, right = self._t._auto_broadcast(a, b)
        return left._derived(trees.Binary("max", left.closure, right.closure))

    def join(self, tensors, dimension=None):
        """(Tensors.scala:560-598)"""
        parts = list(tensors)
        if not parts:
            raise IllegalArgumentException("join of nothing")
        joined = _JoinTensor(self._t, parts)
        if dimension is None:
            return joined
        rank = len(joined.shape)
        if dimension == rank - 1:
            return joined
        dims = [i if i < dimension else (rank - 1 if i == dimension else i - 1) for i in range(rank)]
        return joined.permute(dims)

    def matrixMultiply(self, a, b, variant=capi.GEMM_FP32):
        """C = A . B on MFMA (the reference builds it from split/join, README.md:301-362)."""
        return _GemmTensor(self._t, a, b, variant)


def _flatten(elements):
    """TensorBuilder (Tensors.scala:44-94): nested sequences to (flat data, shape); ragged
    nesting raises IllegalArgumentException."""
    if isinstance(elements, np.ndarray):
        return elements.astype(np.float32).ravel(), tuple(elements.shape)
    if isinstance(elements, (int, float, np.floating, np.integer)):
        return np.array([elements], dtype=np.float32), ()

    def shape_of(x):
        if isinstance(x, np.ndarray):
            return tuple(x.shape)
        if isinstance(x, (list, tuple, Sequence)) and not isinstance(x, (str, bytes)):
            if len(x) == 0:
                return (0,)
            inner = [shape_of(e) for e in x]
            first = inner[0]
            for s in inner[1:]:
                if s != first:
                    raise IllegalArgumentException("ragged nested sequence")
            return (len(x),) + first
        return ()

    shape = shape_of(elements)
    flat = np.asarray(elements, dtype=np.float32).ravel() if _product(shape) else np.zeros(0, np.float32)
    return flat, shape


class Tensors(object):
    """One back-end instance: a device, a stream and the kernel cache in the library.
    Mirror of `trait Tensors` plus the plug-in wiring of `object gpu` (gpu.scala:15-27)."""

    def __init__(self, ordinal=0, math_mode=capi.MATH_PRECISE, vector_width=0, items_per_thread=0,
                 block_size=0, offline=False, axis_reductions=True, fused_sums=True, gemm_idioms="large",
                 join_split=1, launch_order=capi.ORDER_AUTO):
        # launch_order: ORDER_AUTO tiles a kernel's outputs across the dimension along which a
        # transposed leaf's rows run (same bits, fewer bytes fetched); ORDER_LINEAR never does
        self.options = capi.CsFusedOptions(math_mode, vector_width, items_per_thread, block_size, 0, launch_order)
        # The reference's two matrix-multiply idioms (README.md:301-362) as one FP32 MFMA GEMM:
        # "large" when the unrolled kernel would hold at least MM1_MIN_TERMS products per
        # output (matrixMultiply1) or the [i,j,k] product at least MM2_MIN_ELEMENTS elements
        # (matrixMultiply2); "always"; or "never" (the reference's kernels, in its order)
        if gemm_idioms not in ("large", "always", "never"):
            raise IllegalArgumentException("gemm_idioms is one of large, always, never")
        self.gemm_idioms = gemm_idioms
        # How many kernels a join is cut into (a fusion split): 1 (one kernel, the reference's
        # method), a fixed count, or "model" (the learned cost model's choice among the divisors
        # of the part count up to 8, compute/costmodel.py)
        self.join_split = join_split
        self._split_model = None
        # split(d).reduce(+ or max) folds run as one axis reduction (compensated FP32 sums
        # combined in FP64, exact maxima) instead of one kernel with a term per slice
        self.axis_reductions = axis_reductions
        # `sum` of an inline tensor runs as one kernel that computes and adds its elements,
        # without writing them (the reference materialises the tensor first)
        self.fused_sums = fused_sums
        self.Tensor = TensorCompanion(self)
        self._lock = threading.RLock()
        self.launches = 0
        self.idiom_gemms = 0  # matrix-multiply idioms run as one GEMM call
        self.offline = offline
        self._device = self._stream = None
        if offline:
            # Builds closures and generates kernel text without a device (no slow actions).
            return
        dev = ctypes.c_void_p()
        check(lib.cs_device_open(ordinal, ctypes.byref(dev)), "cs_device_open")
        self._device = dev.value
        s = ctypes.c_void_p()
        check(lib.cs_stream_create(self._device, ctypes.byref(s)), "cs_stream_create")
        self._stream = s.value

    # ------------------------------------------------------------------ device info
    def info(self, key):
        v = ctypes.c_int64()
        check(lib.cs_device_info(self._device, key, ctypes.byref(v)), "cs_device_info")
        return v.value

    @property
    def maxComputeUnits(self):
        return self.info(capi.INFO_COMPUTE_UNITS)

    @property
    def stream(self):
        return self._stream

    @property
    def device(self):
        return self._device

    def synchronize(self):
        check(lib.cs_stream_synchronize(self._stream), "cs_stream_synchronize")

    def close(self):
        """monadicClose: drop the kernel cache, then the stream (Tensors.scala:1282-1289)."""
        if self._stream:
            check(lib.cs_fused_cache_clear(self._device), "cs_fused_cache_clear")
            check(lib.cs_stream_destroy(self._stream), "cs_stream_destroy")
            self._stream = None
            check(lib.cs_device_close(self._device), "cs_device_close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ internals
    def _alloc(self, nbytes):
        """A buffer for work on this back-end's stream: stream-ordered, so the host does not
        wait for the device to finish with blocks released earlier."""
        h = ctypes.c_void_p()
        check(lib.cs_buffer_alloc_on(self._stream, max(int(nbytes), 1), ctypes.byref(h)), "cs_buffer_alloc_on")
        return PendingBuffer(self, h.value, int(nbytes))

    def _auto_broadcast(self, a, b):
        """autoBroadcast / autoBroadcastShape (Tensors.scala:202-222): leading dimensions
        align."""
        s1, s2 = a.shape, b.shape
        shape = []
        for i in range(max(len(s1), len(s2))):
            if i >= len(s1) or s1[i] == 1:
                shape.append(s2[i])
            elif i >= len(s2) or s2[i] == 1:
                shape.append(s1[i])
            elif s1[i] == s2[i]:
                shape.append(s1[i])
            else:
                raise IllegalArgumentException(
                    "Failed to automatically broadcast between shape [%s] and [%s]" %
                    (",".join(map(str, s1)), ",".join(map(str, s2))))
        shape = tuple(shape)
        return a.broadcast(shape), b.broadcast(shape)

    def compile(self, closure, shape):
        """Compiles (or fetches) the fused kernel of `closure` over `shape`; returns the
        cs_fused handle (caller releases) and the leaves in argument order."""
        nodes, outputs, leaves = trees.serialize(closure)
        outs = (ctypes.c_int32 * len(outputs))(*outputs)
        shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
        f = ctypes.c_void_p()
        check(lib.cs_fused_compile(self._device, nodes, len(nodes), outs, len(outputs), len(shape), shp,
                                   ctypes.byref(self.options), ctypes.byref(f)), "cs_fused_compile")
        return f.value, leaves, len(outputs)

    def compile_sum(self, closure, shape):
        """The fused-sum kernel of a single-output `closure` over `shape` (cs_fused_compile_sum);
        returns the handle (caller releases) and the leaves in argument order."""
        nodes, outputs, leaves = trees.serialize(closure)
        outs = (ctypes.c_int32 * len(outputs))(*outputs)
        shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
        f = ctypes.c_void_p()
        check(lib.cs_fused_compile_sum(self._device, nodes, len(nodes), outs, len(outputs), len(shape), shp,
                                       ctypes.byref(self.options), ctypes.byref(f)), "cs_fused_compile_sum")
        return f.value, leaves

    def kernel_source(self, closure, shape):
        if self.offline:
            nodes, outputs, _ = trees.serialize(closure)
            outs = (ctypes.c_int32 * len(outputs))(*outputs)
            shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
            src = ctypes.c_void_p()
            check(lib.cs_fused_generate_source(nodes, len(nodes), outs, len(outputs), len(shape), shp,
                                               ctypes.byref(self.options), ctypes.byref(src), None),
                  "cs_fused_generate_source")
            return capi.take_string(src.value)
        f, _, _ = self.compile(closure, shape)
        try:
            src = ctypes.c_char_p()
            check(lib.cs_fused_source(f, ctypes.byref(src)), "cs_fused_source")
            return src.value.decode()
        finally:
            lib.cs_fused_release(f)

    def kernel_cache_contains(self, closure, shape):
        """kernelCache.getIfPresent(closure) != null (Tensors.scala:1293)."""
        nodes, outputs, _ = trees.serialize(closure)
        outs = (ctypes.c_int32 * len(outputs))(*outputs)
        shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
        hit = ctypes.c_int()
        check(lib.cs_fused_lookup(self._device, nodes, len(nodes), outs, len(outputs), len(shape), shp,
                                  ctypes.byref(self.options), ctypes.byref(hit)), "cs_fused_lookup")
        return bool(hit.value)

    def _join_groups(self, join):
        parts = len(join._parts)
        if self.join_split == 1 or parts < 2:
            return 1
        if self.join_split == "model":
            from . import costmodel
            if self._split_model is None:
                self._split_model = costmodel.load_split_model(self)
            if self._split_model is None:
                return 1
            return self._split_model.choose_groups(join)
        g = int(self.join_split)
        if g < 1 or parts % g:
            raise IllegalArgumentException("join_split %d does not divide %d parts" % (g, parts))
        return g

    MM1_MIN_TERMS = 1024
    MM2_MIN_ELEMENTS = 1 << 22

    def _enqueue_closure(self, closure, shape, memo):
        """enqueueClosure (Tensors.scala:1291-1392): compile or fetch, materialise the leaves,
        allocate the output, launch."""
        if self.gemm_idioms != "never":
            mm = _match_mm1(closure, shape)
            if mm is not None and (self.gemm_idioms == "always" or mm[3] * mm[4] >= self.MM1_MIN_TERMS):
                return self._idiom_gemm(mm, memo)
        if self.axis_reductions or self.gemm_idioms != "never":
            fold = _split_fold(closure, shape)
            if fold is not None and fold[1] == 1 and fold[2] == "+" and self.gemm_idioms != "never":
                mm = _match_mm2(fold[0].id)
                if mm is not None and (self.gemm_idioms == "always" or mm[2] * mm[3] * mm[4] >= self.MM2_MIN_ELEMENTS):
                    return self._idiom_gemm(mm, memo)
            if fold is not None and self.axis_reductions:
                out = self._axis_reduce(fold, memo)
                if out is not None:
                    return out
        f, leaves, nout = self.compile(closure, shape)
        try:
            args = [leaf.id._buffer(memo) for leaf in leaves]
            try:
                out = self._alloc(max(_product(shape), 1) * nout * 4)
                ptrs = (ctypes.c_void_p * max(len(args), 1))(*[a.ptr for a in args])
                check(lib.cs_fused_launch(self._stream, f, ptrs, len(args), out.ptr), "cs_fused_launch")
                self.launches += 1
                return out
            finally:
                for a in args:
                    a.release()
        finally:
            lib.cs_fused_release(f)


    def _idiom_gemm(self, mm, memo):
        """C = A . B for a recognised matrix-multiply idiom: one FP32 MFMA GEMM over the
        operands' buffers (B given as [j,k] or as its [1,j,k] reshape: the same memory). The
        [i,j,k] product of matrixMultiply2 is never materialised."""
        a_t, b_t, i, j, k = mm
        a = a_t._buffer(memo)
        try:
            b = b_t._buffer(memo)
            try:
                out = self._alloc(max(i * k, 1) * 4)
                check(lib.cs_gemm(self._stream, capi.GEMM_FP32, i, k, j, a.ptr, j, b.ptr, k, out.ptr, k), "cs_gemm")
                self.launches += 1
                self.idiom_gemms += 1
                return out
            finally:
                b.release()
        finally:
            a.release()

    def _axis_reduce(self, fold, memo):
        """A recognised split fold over dimension d of a dense [outer, n, inner] buffer: row
        kernels when d is innermost, column sums otherwise (max only along the innermost
        dimension, up to 64 column passes). None: not covered, the fold is compiled."""
        param, d, op = fold
        outer = _product(param.shape[:d])
        n = param.shape[d]
        inner = _product(param.shape[d + 1:])
        if inner != 1 and (op == "max" or outer > 64):
            return None
        src = param.id._buffer(memo)
        try:
            out = self._alloc(max(outer * inner, 1) * 4)
            if outer * inner == 0:
                return out
            if inner == 1:
                sums, maxima = (out.ptr, None) if op == "+" else (None, out.ptr)
                check(lib.cs_reduce_rows(self._stream, src.ptr, outer, n, sums, maxima, None, None),
                      "cs_reduce_rows")
            else:
                for o in range(outer):
                    check(lib.cs_reduce_cols(self._stream, src.ptr + o * n * inner * 4, n, inner,
                                             out.ptr + o * inner * 4), "cs_reduce_cols")
            self.launches += 1
            return out
        finally:
            src.release()


_gpu = None
_gpu_lock = threading.Lock()


def gpu():
    """The default back-end on GPU 0 (`object gpu`, gpu.scala:15-27)."""
    global _gpu
    with _gpu_lock:
        if _gpu is None:
            _gpu = Tensors(0)
        return _gpu
