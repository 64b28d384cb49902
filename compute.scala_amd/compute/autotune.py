"""Exhaustive autotuner for fused-expression kernels.

The reference launches every generated kernel with the runtime's default work-group size
and one work item per output (Tensors.scala:1346-1376); it has no tuning. Here each fused
kernel has launch parameters (vector width V, vectors per thread U, workgroup size, and
non-temporal stores) that change its speed but never its results. `tune_fused` times every
candidate on the device (HIP events over back-to-back launches on rotating buffers) and
returns the measurements; `TimingDB` keeps them keyed by the kernel's structural key, so a
later run (or the cost model) can reuse them.
"""
import ctypes
import itertools
import json
import os
import threading

from . import _capi as capi
from ._capi import check, lib

VECTOR_WIDTHS = (1, 2, 4, 8)
ITEMS = (1, 2, 4)
BLOCKS = (64, 128, 256, 512, 1024)
NONTEMPORAL = (0, 1)


def candidates(last_dim, vector_widths=VECTOR_WIDTHS, items=ITEMS, blocks=BLOCKS, nontemporal=NONTEMPORAL):
    """Every launch configuration valid for an output whose innermost dimension is
    `last_dim` (V must divide it)."""
    out = []
    for v, u, b, nt in itertools.product(vector_widths, items, blocks, nontemporal):
        if last_dim % v:
            continue
        out.append({"vector_width": v, "items_per_thread": u, "block_size": b, "nontemporal_store": nt})
    return out


def options_struct(math_mode, cfg):
    return capi.CsFusedOptions(math_mode, cfg["vector_width"], cfg["items_per_thread"], cfg["block_size"],
                               cfg.get("nontemporal_store", 0))


def time_launches(tensors, fused, leaf_sets, outs, reps):
    """Average microseconds per launch of `fused`, cycling through the given buffer sets,
    replayed from a captured graph and timed with HIP events on the launch stream."""
    stream = tensors.stream
    n = len(outs)
    ptrs = [(ctypes.c_void_p * max(len(ls), 1))(*ls) for ls in leaf_sets]
    nleaves = len(leaf_sets[0])
    for i in range(n):
        check(lib.cs_fused_launch(stream, fused, ptrs[i], nleaves, outs[i]))
    check(lib.cs_graph_begin_capture(stream))
    try:
        for i in range(n):
            check(lib.cs_fused_launch(stream, fused, ptrs[i], nleaves, outs[i]))
    finally:
        g = ctypes.c_void_p()
        check(lib.cs_graph_end_capture(stream, ctypes.byref(g)))
    try:
        check(lib.cs_graph_launch(stream, g))
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib.cs_stream_synchronize(stream))
        check(lib.cs_event_record(stream, ctypes.byref(e0)))
        for _ in range(reps):
            check(lib.cs_graph_launch(stream, g))
        check(lib.cs_event_record(stream, ctypes.byref(e1)))
        check(lib.cs_event_wait(e1))
        ms = ctypes.c_float()
        check(lib.cs_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        lib.cs_event_release(e0)
        lib.cs_event_release(e1)
    finally:
        lib.cs_graph_release(g)
    return ms.value * 1e3 / (reps * n)


def compile_with(tensors, closure, shape, cfg):
    """cs_fused handle of `closure` compiled with launch configuration `cfg`."""
    from . import trees
    nodes, outputs, leaves = trees.serialize(closure)
    outs = (ctypes.c_int32 * len(outputs))(*outputs)
    shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
    opts = options_struct(tensors.options.math_mode, cfg)
    f = ctypes.c_void_p()
    check(lib.cs_fused_compile(tensors.device, nodes, len(nodes), outs, len(outputs), len(shape), shp,
                               ctypes.byref(opts), ctypes.byref(f)), "cs_fused_compile")
    return f.value, leaves


def tune_fused(tensors, closure, shape, leaf_sets, outs, configs, reps=8):
    """Times every configuration; returns [(cfg, microseconds)] sorted fastest first."""
    results = []
    for cfg in configs:
        f, _ = compile_with(tensors, closure, shape, cfg)
        try:
            us = time_launches(tensors, f, leaf_sets, outs, reps)
        finally:
            lib.cs_fused_release(f)
        results.append((cfg, us))
    results.sort(key=lambda r: r[1])
    return results


def structural_key(tensors, closure, shape):
    from . import trees
    nodes, outputs, _ = trees.serialize(closure)
    outs = (ctypes.c_int32 * len(outputs))(*outputs)
    shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
    src = ctypes.c_void_p()
    key = ctypes.c_uint64()
    check(lib.cs_fused_generate_source(nodes, len(nodes), outs, len(outputs), len(shape), shp, None,
                                       ctypes.byref(src), ctypes.byref(key)))
    capi.take_string(src.value)
    return "%016x" % key.value


class TimingDB(object):
    """Measured launch times keyed by (structural key, configuration), stored as JSON."""

    def __init__(self, path=None):
        self.path = path
        self._lock = threading.Lock()
        self.rows = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.rows = json.load(f)

    @staticmethod
    def cfg_key(cfg):
        return "v%d_u%d_b%d_nt%d" % (cfg["vector_width"], cfg["items_per_thread"], cfg["block_size"],
                                     cfg.get("nontemporal_store", 0))

    def record(self, key, cfg, us):
        with self._lock:
            self.rows.setdefault(key, {})[self.cfg_key(cfg)] = us

    def best(self, key):
        entries = self.rows.get(key)
        if not entries:
            return None
        name = min(entries, key=lambda k: (entries[k], k))
        return name, entries[name]

    def save(self):
        if not self.path:
            return
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.rows, f, indent=0, sort_keys=True)
        os.replace(tmp, self.path)


# ------------------------------------------------------------------------ fusion splits
SPLIT_CONFIG = {"vector_width": 1, "items_per_thread": 1, "block_size": 256, "nontemporal_store": 0}


def split_candidates(parts, largest=8):
    """The group counts a join of `parts` parts can be cut into: divisors up to `largest`."""
    return [g for g in (1, 2, 4, 8, 16) if g <= largest and parts % g == 0]


def time_join(tensors, join, leaf_ptrs, groups, cfg=SPLIT_CONFIG, reps=5):
    """Microseconds to materialise the join tensor `join` (a Tensor.join) cut into `groups`
    kernels plus the gathering kernel (groups > 1), replayed from a captured graph of those
    launches and timed with HIP events. leaf_ptrs maps id(leaf tensor) to its device pointer."""
    from . import tensors as tz
    from . import trees
    heads, per = join.closures(groups)
    head = tuple(join._head_shape)
    n_head = 1
    for s in head:
        n_head *= s
    plan, keep = [], []
    try:
        if groups == 1:
            heads = [trees.Concatenate([p.closure for p in join._parts])]
            per = len(join._parts)
        part_ptrs = []
        for h in heads:
            f, leaves = compile_with(tensors, h, head, cfg)
            out = tensors._alloc(max(n_head * per, 1) * 4)
            keep.append(out)
            plan.append((f, [leaf_ptrs[id(l.id)] for l in leaves], out.ptr))
            part_ptrs.append(out.ptr)
        if groups > 1:
            shells = [tz.CachedTensor(tensors, head + (per,), 0.0, None) for _ in range(groups)]
            ptr_of = {id(s): p for s, p in zip(shells, part_ptrs)}
            r = len(head)
            gather = trees.Concatenate([trees.Extract(trees.Transform(s.arrayTerm, tz._split_matrix(r + 1, r, j), r))
                                        for s in shells for j in range(per)])
            f, leaves = compile_with(tensors, gather, head, cfg)
            out = tensors._alloc(max(n_head * len(join._parts), 1) * 4)
            keep.append(out)
            plan.append((f, [ptr_of[id(l.id)] for l in leaves], out.ptr))
        arrays = [(ctypes.c_void_p * max(len(p), 1))(*p) for _, p, _ in plan]

        def run():
            for (f, p, o), a in zip(plan, arrays):
                check(lib.cs_fused_launch(tensors.stream, f, a, len(p), o))

        run()
        tensors.synchronize()
        check(lib.cs_graph_begin_capture(tensors.stream))
        try:
            run()
        finally:
            g = ctypes.c_void_p()
            check(lib.cs_graph_end_capture(tensors.stream, ctypes.byref(g)))
        try:
            check(lib.cs_graph_launch(tensors.stream, g))
            e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
            check(lib.cs_stream_synchronize(tensors.stream))
            check(lib.cs_event_record(tensors.stream, ctypes.byref(e0)))
            for _ in range(reps):
                check(lib.cs_graph_launch(tensors.stream, g))
            check(lib.cs_event_record(tensors.stream, ctypes.byref(e1)))
            check(lib.cs_event_wait(e1))
            ms = ctypes.c_float()
            check(lib.cs_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
            lib.cs_event_release(e0)
            lib.cs_event_release(e1)
        finally:
            lib.cs_graph_release(g)
        return ms.value * 1e3 / reps
    finally:
        for f, _, _ in plan:
            lib.cs_fused_release(f)
        for b in keep:
            b.release()


def split_kernels(join, groups):
    """(closure, shape) of every kernel a split of `join` into `groups` compiles, the gathering
    kernel's leaves standing in as shells (for precompiling into the code-object cache)."""
    from . import tensors as tz
    from . import trees
    head = tuple(join._head_shape)
    if groups == 1:
        return [(trees.Concatenate([p.closure for p in join._parts]), head)]
    heads, per = join.closures(groups)
    shells = [tz.CachedTensor(join._tensors, head + (per,), 0.0, None) for _ in range(groups)]
    r = len(head)
    gather = trees.Concatenate([trees.Extract(trees.Transform(s.arrayTerm, tz._split_matrix(r + 1, r, j), r))
                                for s in shells for j in range(per)])
    return [(h, head) for h in heads] + [(gather, head)]


def precompile(closure, shape, cfg, math_mode=0):
    """Compiles `closure`'s kernel into the on-disk code-object cache without a device
    (cs_fused_precompile), so that a later compile of it on a device only loads it."""
    from . import trees
    nodes, outputs, _ = trees.serialize(closure)
    outs = (ctypes.c_int32 * len(outputs))(*outputs)
    shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
    opts = options_struct(math_mode, cfg)
    check(lib.cs_fused_precompile(nodes, len(nodes), outs, len(outputs), len(shape), shp, ctypes.byref(opts)),
          "cs_fused_precompile")
