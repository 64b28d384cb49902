"""Many fused expressions on one device: per-expression launches and the persistent batch.

The reference materialises each expression with its own kernel launch (enqueueClosure,
Tensors.scala:1291-1392). `ExpressionSet` holds a set of expressions with their leaves and
outputs resident on the device and can run them either way:

  * `launch_each(configs)`: one launch per expression (the reference's method), each with its
    own vector width V, vectors per thread U and workgroup size;
  * `Batch`: one persistent kernel (cs_batch) that drains a queue of (expression, tile)
    work items, captured in a graph so that one replay runs everything.

Both evaluate the same generated bodies, so their outputs are bit-identical.
"""
import ctypes

import numpy as np

from . import _capi as capi
from . import trees
from ._capi import check, lib

DEFAULT = {"vector_width": 4, "items_per_thread": 1, "block_size": 256}


def _descs(exprs, configs, math_mode):
    """cs_fused_desc for each (expression, configuration); keeps the ctypes buffers alive."""
    keep, descs = [], (capi.CsFusedDesc * len(exprs))()
    for i, (e, cfg) in enumerate(zip(exprs, configs)):
        nodes, outputs, leaves = trees.serialize(e.tensor.closure)
        outs = (ctypes.c_int32 * len(outputs))(*outputs)
        shp = (ctypes.c_int32 * max(len(e.shape), 1))(*e.shape)
        keep += [nodes, outs, shp]
        d = descs[i]
        d.nodes = ctypes.cast(nodes, ctypes.POINTER(capi.CsNode))
        d.nnodes = len(nodes)
        d.outputs = ctypes.cast(outs, ctypes.POINTER(ctypes.c_int32))
        d.noutputs = len(outputs)
        d.out_rank = len(e.shape)
        d.out_shape = ctypes.cast(shp, ctypes.POINTER(ctypes.c_int32))
        v = cfg["vector_width"]
        last = e.shape[-1] if e.shape else 1
        while v > 1 and last % v:
            v //= 2
        d.options = capi.CsFusedOptions(math_mode, v, cfg["items_per_thread"], cfg.get("block_size", 256),
                                        cfg.get("nontemporal_store", 0))
    return descs, keep


def compile_many(tensors, exprs, configs):
    """cs_fused handles, one hiprtc module for all of them."""
    descs, keep = _descs(exprs, configs, tensors.options.math_mode)
    out = (ctypes.c_void_p * len(exprs))()
    check(lib.cs_fused_compile_many(tensors.device, descs, len(exprs), out), "cs_fused_compile_many")
    return [out[i] for i in range(len(exprs))]


def effective_vector_width(shape, v):
    last = shape[-1] if shape else 1
    while v > 1 and last % v:
        v //= 2
    return v


class ExpressionSet(object):
    """Expressions with device-resident leaves and outputs."""

    def __init__(self, tensors, exprs):
        self.T = tensors
        self.exprs = list(exprs)
        self.leaf_bufs = []
        self.out_bufs = []
        for e in self.exprs:
            bufs = []
            for arr in e.leaf_data:
                a = np.ascontiguousarray(arr, dtype=np.float32).ravel()
                b = tensors._alloc(max(a.size, 1) * 4)
                check(lib.cs_buffer_write(tensors.stream, b.handle, 0, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 0))
                bufs.append(b)
            self.leaf_bufs.append(bufs)
            self.out_bufs.append(tensors._alloc(max(e.out_elements, 1) * 4))
        tensors.synchronize()
        self._ptrs = [(ctypes.c_void_p * max(len(b), 1))(*[x.ptr for x in b]) for b in self.leaf_bufs]

    @property
    def bytes(self):
        return sum(e.bytes for e in self.exprs)

    def launch(self, i, fused, block=0):
        check(lib.cs_fused_launch_block(self.T.stream, fused, self._ptrs[i], len(self.leaf_bufs[i]),
                                        self.out_bufs[i].ptr, block))

    def outputs(self):
        res = []
        for e, b in zip(self.exprs, self.out_bufs):
            res.append(b.to_numpy(e.out_elements))
        return res

    def clear_outputs(self):
        for b in self.out_bufs:
            check(lib.cs_fill_uniform(self.T.stream, b.ptr, b.nbytes // 4, 7, -9.0, 9.0))

    def release(self):
        for bufs in self.leaf_bufs:
            for b in bufs:
                b.release()
        for b in self.out_bufs:
            b.release()
        self.leaf_bufs, self.out_bufs = [], []


class Graph(object):
    """A captured sequence of launches on the set's stream, timed with HIP events."""

    def __init__(self, tensors, fn):
        self.T = tensors
        fn()  # eager run: sizes scratch, loads modules
        tensors.synchronize()
        check(lib.cs_graph_begin_capture(tensors.stream))
        try:
            fn()
        finally:
            g = ctypes.c_void_p()
            check(lib.cs_graph_end_capture(tensors.stream, ctypes.byref(g)))
        self.handle = g

    def launch(self):
        check(lib.cs_graph_launch(self.T.stream, self.handle))

    def time_us(self, reps=3):
        """Average microseconds of one replay."""
        s = self.T.stream
        self.launch()
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib.cs_stream_synchronize(s))
        check(lib.cs_event_record(s, ctypes.byref(e0)))
        for _ in range(reps):
            self.launch()
        check(lib.cs_event_record(s, ctypes.byref(e1)))
        check(lib.cs_event_wait(e1))
        ms = ctypes.c_float()
        check(lib.cs_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        lib.cs_event_release(e0)
        lib.cs_event_release(e1)
        return ms.value * 1e3 / reps

    def release(self):
        if self.handle:
            lib.cs_graph_release(self.handle)
            self.handle = None


class Batch(object):
    """One persistent kernel over every expression of an ExpressionSet (cs_batch)."""

    def __init__(self, eset, fused, block=256):
        self.eset = eset
        arr = (ctypes.c_void_p * len(fused))(*fused)
        h = ctypes.c_void_p()
        check(lib.cs_batch_create(eset.T.device, arr, len(fused), block, ctypes.byref(h)), "cs_batch_create")
        self.handle = h.value
        args = []
        for i in range(len(eset.exprs)):
            args += [b.ptr for b in eset.leaf_bufs[i]] + [eset.out_bufs[i].ptr]
        table = (ctypes.c_void_p * len(args))(*args)
        check(lib.cs_batch_bind(eset.T.stream, self.handle, table, len(args)), "cs_batch_bind")

    def launch(self):
        check(lib.cs_batch_launch(self.eset.T.stream, self.handle))

    def info(self):
        items, grid, block = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        check(lib.cs_batch_info(self.handle, ctypes.byref(items), ctypes.byref(grid), ctypes.byref(block)))
        return {"items": items.value, "grid": grid.value, "block": block.value}

    def release(self):
        if self.handle:
            lib.cs_batch_release(self.handle)
            self.handle = None
