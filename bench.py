"""Benchmark of the fused-expression back-end on one MI355X.

Headline (the `value` of the JSON line): the view-heavy fused chain
    X [64,256,256] ~ N(0,1), bias [256] ~ U(-0.5,0.5)
    p = X.permute(1,0,2); y = p + bias.broadcast([256,64,256]) + p.translate([0,0,1])
    out = exp(-(y*y)) * tanh(y)
as one generated kernel, in GB/s of algorithmic traffic (X read once, bias, out written:
33,555,456 bytes per launch). One step = one launch over one input set; 16 input/output
sets rotate (512 MiB in total, twice the 256 MB Infinity Cache) so every launch streams
from HBM. The K timed launches are spread over 3 streams (launch i on stream
(i mod 16) mod 3, as the reference's CommandQueuePool leases a queue per command and its JMH
benchmarks run independent iterations from every core at once, benchmarks.scala:31-32, 56),
each stream replaying captured hipGraphs of its share, and timed from an event recorded on
the first stream before any launch is submitted to the latest end event of the streams.
Beside it: the same K launches on one stream, and a window
of at least 4096 launches over 1 to 4 streams.

The same line carries the other workloads under "also": a*b+c on [1024,1024] and its sum,
the one-pass row sum/row max/total over [16384,4096], and C = A.B at 4096^3 per GEMM
variant with its FP64 error. `cpu_baseline` times the reference's generated kernel for the
headline workload (its OpenCL C text run as host C on OpenMP threads) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only c2,c1,c3,c4,conv,c5]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "compute.scala_amd"))
sys.path.insert(0, ROOT)

METRIC = "fused-expr HBM GB/s (% of 8 TB/s); GEMM TFLOP/s at err≤1e-5 vs FP64; vs plain"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0, "fp8": 5000.0}
# Input/output sets the headline rotates through: 16 x 32 MiB = 512 MiB, twice the 256 MB
# Infinity Cache, and within the GPU's TLB reach (past 16 sets the headline kernel, whose
# permuted read touches every 2 MiB page of X from the first wave, slows by a fifth).
HEADLINE_SETS = 16


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


class Timer(object):
    """HIP events on the launch stream."""

    def __init__(self, capi, stream):
        self.capi, self.stream = capi, stream

    def measure(self, fn, reps):
        lib = self.capi.lib
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        self.capi.check(lib.cs_stream_synchronize(self.stream))
        self.capi.check(lib.cs_event_record(self.stream, ctypes.byref(e0)))
        for _ in range(reps):
            fn()
        self.capi.check(lib.cs_event_record(self.stream, ctypes.byref(e1)))
        self.capi.check(lib.cs_event_wait(e1))
        self.capi.check(lib.cs_stream_synchronize(self.stream))
        ms = ctypes.c_float()
        self.capi.check(lib.cs_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        lib.cs_event_release(e0)
        lib.cs_event_release(e1)
        return ms.value


def kernel_us_timed(capi, stream, f, ptrs, outs, count):
    """Average duration of `count` back-to-back launches of the fused kernel `f` (cycling over
    the leaf sets `ptrs`), each timed by the events its own dispatch carries
    (cs_fused_launch_timed): the kernel's duration without the gap before it, the figure a
    profiler reports per dispatch."""
    lib = capi.lib
    capi.check(lib.cs_stream_synchronize(stream))
    evs = []
    for i in range(count):
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        k = i % len(ptrs)
        capi.check(lib.cs_fused_launch_timed(stream, f, ptrs[k], len(ptrs[k]), outs[k], ctypes.byref(e0),
                                             ctypes.byref(e1)))
        evs.append((e0, e1))
    capi.check(lib.cs_stream_synchronize(stream))
    total = 0.0
    for e0, e1 in evs:
        ms = ctypes.c_float()
        capi.check(lib.cs_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        total += ms.value
        lib.cs_event_release(e0)
        lib.cs_event_release(e1)
    return total * 1e3 / count


def capture(capi, stream, fns):
    lib = capi.lib
    capi.check(lib.cs_graph_begin_capture(stream))
    for f in fns:
        f()
    g = ctypes.c_void_p()
    capi.check(lib.cs_graph_end_capture(stream, ctypes.byref(g)))
    return g


class StreamSpread(object):
    """Independent launches spread over `k` streams, the launch stream first. Launch i works on
    set s = i mod `sets` and runs on stream s mod k, so each set stays on one stream and no
    two streams touch the same buffers (when k divides `sets` the streams share the launches
    evenly). This is the reference's CommandQueuePool, which leases one of its queues to each
    command (OpenCL.scala:423-448, 1313-1322; five queues per device in gpu.scala:26). Each
    stream replays captured graphs of its own launches, so one launch's ramp-down overlaps
    another stream's work. A run is timed from an event on the first stream, recorded before
    any launch is submitted, to the latest end event of the streams, which are idle before it
    and synchronised after it. With k = 1 it is the launch stream alone."""

    ROUNDS = 16  # rounds of the sets in one replay of the full graphs

    def __init__(self, T, capi, launch_on, sets, k, prime=False):
        self.T, self.capi, self.lib = T, capi, capi.lib
        self.prime = prime  # run each stream's sets once before capturing its graphs
        self.launch_on, self.sets, self.k = launch_on, sets, k
        self.streams, self.made = [T.stream], []
        for _ in range(k - 1):
            st = ctypes.c_void_p()
            capi.check(self.lib.cs_stream_create(T.device, ctypes.byref(st)))
            self.streams.append(st)
            self.made.append(st)
        self.block = sets * self.ROUNDS
        self.full = self._graphs(self.block)
        self.tails = {}

    def _graphs(self, count):
        """One graph per stream holding its share of launches 0 .. count-1."""
        gs = []
        for j, st in enumerate(self.streams):
            mine = [i % self.sets for i in range(count) if (i % self.sets) % self.k == j]
            if mine and self.prime:
                # sizes the stream's scratch (reductions) outside the capture
                for i in sorted(set(mine)):
                    self.launch_on(st, i)
                self.capi.check(self.lib.cs_stream_synchronize(st))
            gs.append(capture(self.capi, st, [lambda i=i, st=st: self.launch_on(st, i) for i in mine])
                      if mine else None)
        return gs

    def prepare(self, count):
        """Captures the graphs a run of `count` launches needs, without running them."""
        rem = count % self.block
        if rem and rem not in self.tails:
            self.tails[rem] = self._graphs(rem)

    def run(self, count):
        """Exactly `count` launches (whole replays of the full graphs, then the graphs of the
        remaining count mod block launches); returns the milliseconds they took."""
        lib, check = self.lib, self.capi.check
        self.prepare(count)
        rem = count % self.block
        for st in self.streams:
            check(lib.cs_stream_synchronize(st))
        # The window opens with one event on the first stream, recorded before any graph is
        # submitted: work on the other streams is submitted after it, so it cannot start
        # earlier, and no further event sits between the submissions. It closes with the last
        # of the streams' end events.
        starts, ends, used = [], [], []

        def launch(j, st, g):
            if not used:
                e = ctypes.c_void_p()
                check(lib.cs_event_record(st, ctypes.byref(e)))
                starts.append(e)
            if j not in used:
                used.append(j)
            check(lib.cs_graph_launch(st, g))

        for _ in range(count // self.block):
            for j, (st, g) in enumerate(zip(self.streams, self.full)):
                launch(j, st, g)
        if rem:
            for j, (st, g) in enumerate(zip(self.streams, self.tails[rem])):
                if g is not None:
                    launch(j, st, g)
        for j in used:
            e = ctypes.c_void_p()
            check(lib.cs_event_record(self.streams[j], ctypes.byref(e)))
            ends.append(e)
        for st in self.streams:
            check(lib.cs_stream_synchronize(st))
        first, last = 0.0, None
        for e1 in ends:
            b = ctypes.c_float()
            check(lib.cs_event_elapsed_ms(starts[0], e1, ctypes.byref(b)))
            last = b.value if last is None else max(last, b.value)
        for e in starts + ends:
            lib.cs_event_release(e)
        return last - first

    def release(self):
        for gs in [self.full] + list(self.tails.values()):
            for g in gs:
                if g is not None:
                    self.lib.cs_graph_release(g)
        for st in self.made:
            self.capi.check(self.lib.cs_stream_synchronize(st))
            self.lib.cs_stream_destroy(st)


def bench_c2(T, capi, steps, warmup, sets=HEADLINE_SETS, tune=True, fixed=None, streams=3):
    from compute import autotune
    from compute import tensors as tz
    lib = capi.lib
    n = 64 * 256 * 256
    xs = [T._alloc(n * 4) for _ in range(sets)]
    bs = [T._alloc(256 * 4) for _ in range(sets)]
    outs = [T._alloc(n * 4) for _ in range(sets)]
    for i in range(sets):
        capi.check(lib.cs_fill_normal(T.stream, xs[i].ptr, n, 2001 + 97 * i))
        capi.check(lib.cs_fill_uniform(T.stream, bs[i].ptr, 256, 2002 + 97 * i, -0.5, 0.5))
    # The kernel is the one the Tensor API generates for the expression.
    X = tz.CachedTensor(T, (64, 256, 256), 0.0, xs[0].retain())
    B = tz.CachedTensor(T, (256,), 0.0, bs[0].retain())
    p = X.permute([1, 0, 2])
    y = p + B.broadcast((256, 64, 256)) + p.translate([0, 0, 1])
    out = T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)
    leaf_sets = [[xs[i].ptr, bs[i].ptr] for i in range(sets)]
    out_ptrs = [o.ptr for o in outs]
    tuning = None
    best_cfg = None
    if fixed is not None:
        best_cfg = fixed
        tuning = {"best": fixed, "fixed": True}
    elif tune:
        cands = autotune.candidates(256, vector_widths=(4, 8), items=(1, 2, 4), blocks=(128, 256, 512, 1024),
                                    nontemporal=(0, 1))
        results = autotune.tune_fused(T, out.closure, out.shape, leaf_sets, out_ptrs, cands, reps=8)
        best_cfg = results[0][0]
        default_us = [us for cfg, us in results if cfg == {"vector_width": 4, "items_per_thread": 2,
                                                          "block_size": 256, "nontemporal_store": 0}]
        tuning = {"best": best_cfg, "best_us": results[0][1], "default_us": default_us[0] if default_us else None,
                  "worst_us": results[-1][1], "candidates": len(results)}
        log("c2 tuning: best %s %.2f us, default %s us, worst %.2f us" %
            (best_cfg, results[0][1], tuning["default_us"], results[-1][1]))
    if best_cfg is not None:
        f, leaves = autotune.compile_with(T, out.closure, out.shape, best_cfg)
    else:
        f, leaves, _ = T.compile(out.closure, out.shape)
    assert [l.id for l in leaves] == [X, B]
    ptrs = [(ctypes.c_void_p * 2)(xs[i].ptr, bs[i].ptr) for i in range(sets)]

    def launch(i):
        capi.check(lib.cs_fused_launch(T.stream, f, ptrs[i], 2, outs[i].ptr))

    def launch_on(st, i):
        capi.check(lib.cs_fused_launch(st, f, ptrs[i], 2, outs[i].ptr))

    for i in range(sets):
        launch(i)
    T.synchronize()
    steps = max(1, steps)
    bytes_per_launch_c2 = 4 * n + 4 * 256 + 4 * n
    # The timed region: exactly `steps` launches after exactly `warmup` launches, spread over
    # `streams` streams (the launch stream and streams - 1 more).
    spread = StreamSpread(T, capi, launch_on, sets, streams)
    spread.prepare(warmup)
    spread.prepare(steps)
    if warmup > 0:
        spread.run(warmup)
    ms = spread.run(steps)
    launches = steps
    # ten more runs of the same K launches (after the timed one): how much the figure moves
    reruns = sorted(bytes_per_launch_c2 * steps / (spread.run(steps) * 1e-3) / 1e9 for _ in range(10))
    # The same kernel over a window of at least 4096 launches (about 20-30 ms), a figure that
    # does not depend on how short the requested run is, spread over 1 to 4 streams.
    long_launches = max(4096, steps)
    ms_long = spread.run(long_launches)
    spread.release()
    by_streams = {}
    ms_one = ms
    for k in (1, 2, 3, 4):
        if k == streams:
            by_streams[k] = round(bytes_per_launch_c2 * long_launches / (ms_long * 1e-3) / 1e9, 1)
            continue
        other = StreamSpread(T, capi, launch_on, sets, k)
        other.run(sets)
        if k == 1:
            ms_one = other.run(steps)
        ms_k = other.run(long_launches)
        by_streams[k] = round(bytes_per_launch_c2 * long_launches / (ms_k * 1e-3) / 1e9, 1)
        other.release()
    timer = Timer(capi, T.stream)
    reps = max(1, long_launches // sets)
    # hot-cache figure: one input set, repeated
    g_hot = capture(capi, T.stream, [lambda: launch(0)] * sets)
    ms_hot = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g_hot)), reps)
    us_event = kernel_us_timed(capi, T.stream, f, ptrs, [o.ptr for o in outs], 16 * sets)
    # Reference point: the runtime's device-to-device copy of X into the output buffer (the
    # same bytes in and out, the same rotating sets and graph timing) -- the streaming rate a
    # kernel with no arithmetic reaches at this size.
    g_copy = capture(capi, T.stream, [lambda i=i: capi.check(lib.cs_memcpy_dtod(T.stream, outs[i].ptr, xs[i].ptr, 4 * n))
                                      for i in range(sets)])
    timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g_copy)), 2)
    ms_copy = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g_copy)), reps)
    lib.cs_graph_release(g_copy)
    bytes_per = 4 * n + 4 * 256 + 4 * n
    res = dict(bytes_per_launch=bytes_per, launches=launches, ms=ms, ms_per_launch=ms / launches,
               us_per_launch_events=us_event,
               one_stream={"launches": steps, "us_per_launch": ms_one * 1e3 / steps,
                           "gbs": bytes_per * steps / (ms_one * 1e-3) / 1e9},
               reruns={"runs": len(reruns), "min_gbs": round(reruns[0], 1),
                       "median_gbs": round((reruns[4] + reruns[5]) / 2, 1), "max_gbs": round(reruns[-1], 1)},
               long_window={"launches": long_launches, "streams": streams,
                            "us_per_launch": ms_long * 1e3 / long_launches,
                            "gbs": bytes_per * long_launches / (ms_long * 1e-3) / 1e9,
                            "gbs_by_streams": by_streams},
               copy_reference={"bytes": 8 * n, "us": ms_copy * 1e3 / (reps * sets),
                               "gbs": 8 * n * reps * sets / (ms_copy * 1e-3) / 1e9},
               gbs=bytes_per * launches / (ms * 1e-3) / 1e9,
               gbs_hot_cache=bytes_per * reps * sets / (ms_hot * 1e-3) / 1e9, tuning=tuning)
    lib.cs_graph_release(g_hot)
    # PCIe-inclusive rate, reported beside `value` and never as it: X and bias handed over as
    # host arrays through the Tensor API's upload path (pageable memory -> pinned staging ->
    # device), the kernel, and the output read back to a host array; wall clock per step.
    hx = np.random.default_rng(7).standard_normal(n).astype(np.float32)
    hb = np.random.default_rng(8).uniform(-0.5, 0.5, 256).astype(np.float32)
    hout = np.empty(n, np.float32)

    def host_step():
        capi.check(lib.cs_buffer_write(T.stream, xs[0].handle, 0, hx.ctypes.data_as(ctypes.c_void_p), 4 * n, 0))
        capi.check(lib.cs_buffer_write(T.stream, bs[0].handle, 0, hb.ctypes.data_as(ctypes.c_void_p), 4 * 256, 0))
        launch(0)
        capi.check(lib.cs_buffer_read(T.stream, outs[0].handle, 0, hout.ctypes.data_as(ctypes.c_void_p), 4 * n))

    # untimed steps first: the first ones after the arrays are created run up to 1.7x slower
    warm = 10
    for _ in range(warm):
        host_step()
    k = 20
    t0 = time.perf_counter()
    for _ in range(k):
        host_step()
    pcie_us = (time.perf_counter() - t0) * 1e6 / k
    res["pcie_inclusive"] = {"us_per_step": round(pcie_us, 1), "gbs": round(bytes_per / (pcie_us * 1e-6) / 1e9, 2),
                             "steps": k, "warmup": warm,
                             "path": "host arrays in and out: cs_buffer_write (pinned staging), "
                                     "the fused kernel, cs_buffer_read; wall clock"}
    log("c2 with host arrays in and out (PCIe included): %.1f us per step, %.2f GB/s" %
        (pcie_us, res["pcie_inclusive"]["gbs"]))
    lib.cs_fused_release(f)
    X.release()
    B.release()
    for h in xs + bs + outs:
        h.release()
    return res


def bench_hbm(T, capi, launches=1024, streams=3):
    """The library's own streaming kernels (cs_hbm_probe: no arithmetic, the generated
    kernels' load/store forms and layout) under the headline's harness -- rotating buffer
    sets, graph replay, HIP events -- at the headline's size (16 MiB read + 16 MiB written per
    launch, the headline's sets) and at 256 MiB per launch: what HBM gives a kernel of that
    size and shape. At 16 MiB every copy form is also spread over `streams` streams over 4096
    launches, as the headline's long window is."""
    lib = capi.lib
    out = {}
    names = {capi.PROBE_STORE_PLAIN: "plain", capi.PROBE_STORE_NT: "nt", capi.PROBE_STORE_NT_WRITE_THROUGH: "sc1 nt"}
    for label, n, sets in (("16MiB", 1 << 22, HEADLINE_SETS), ("256MiB", 1 << 26, 2)):
        srcs = [T._alloc(n * 4) for _ in range(sets)]
        dsts = [T._alloc(n * 4) for _ in range(sets)]
        for i in range(sets):
            capi.check(lib.cs_fill_uniform(T.stream, srcs[i].ptr, n, 7000 + i, 0.0, 1.0))
        T.synchronize()
        rows = []
        combos = [(capi.PROBE_READ, capi.PROBE_STORE_PLAIN, u) for u in (1, 4)]
        combos += [(k, st, u) for k in (capi.PROBE_WRITE, capi.PROBE_COPY) for st in names for u in (1, 2, 4)]
        timer = Timer(capi, T.stream)

        def one_on(st, i, kind, store, items):
            capi.check(lib.cs_hbm_probe(st, kind, store, srcs[i].ptr, dsts[i].ptr, n, items))

        for kind, store, items in combos:
            def one(i, kind=kind, store=store, items=items):
                one_on(T.stream, i, kind, store, items)
            for i in range(sets):
                one(i)
            g = capture(capi, T.stream, [lambda i=i: one(i) for i in range(sets)])
            reps = max(1, launches // sets) if sets > 2 else 16
            timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g)), 2)
            ms = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g)), reps)
            lib.cs_graph_release(g)
            nbytes = 4 * n * (2 if kind == capi.PROBE_COPY else 1)
            us = ms * 1e3 / (reps * sets)
            rows.append({"kind": ["read", "write", "copy"][kind], "store": names[store] if kind else "-",
                         "items_per_thread": items, "bytes": nbytes, "us": round(us, 3),
                         "gbs": round(nbytes / (us * 1e-6) / 1e9, 1)})
        best = {k: max((r for r in rows if r["kind"] == k), key=lambda r: r["gbs"]) for k in ("read", "write", "copy")}
        out[label] = {"best": best, "all": rows}
        if label == "16MiB":
            # every copy form spread over the streams, as the headline is: its ceiling
            spread_rows = []
            for store in names:
                for items in (1, 2, 4):
                    spread = StreamSpread(T, capi, lambda st, i, store=store, items=items:
                                          one_on(st, i, capi.PROBE_COPY, store, items), sets, streams)
                    spread.run(sets)
                    count = 4096
                    ms = spread.run(count)
                    spread.release()
                    spread_rows.append({"streams": streams, "launches": count, "store": names[store],
                                        "items_per_thread": items, "us_per_launch": round(ms * 1e3 / count, 3),
                                        "gbs": round(8 * n * count / (ms * 1e-3) / 1e9, 1)})
            out[label]["copy_spread"] = max(spread_rows, key=lambda r: r["gbs"])
            out[label]["copy_spread_all"] = spread_rows
        for h in srcs + dsts:
            h.release()
        log("hbm probe %s per array: best read %.0f, write %.0f, copy %.0f GB/s%s" %
            (label, best["read"]["gbs"], best["write"]["gbs"], best["copy"]["gbs"],
             ", copy over %d streams %.0f GB/s" % (streams, out[label]["copy_spread"]["gbs"])
             if "copy_spread" in out[label] else ""))
    return out


def bench_c1(T, capi, steps, sets=33, streams=3):
    lib = capi.lib
    n = 1024 * 1024
    a = [[T._alloc(n * 4) for _ in range(3)] for _ in range(sets)]
    outs = [T._alloc(n * 4) for _ in range(sets)]
    sums = [T._alloc(4) for _ in range(sets)]
    for i in range(sets):
        for j in range(3):
            capi.check(lib.cs_fill_uniform(T.stream, a[i][j].ptr, n, 1001 + j + 97 * i, -1.0, 1.0))
    from compute import tensors as tz
    ta, tb, tc = (tz.CachedTensor(T, (1024, 1024), 0.0, a[0][j].retain()) for j in range(3))
    e = ta * tb + tc
    # launch configuration from the same autotuner sweep as the headline kernel
    from compute import autotune
    cands = autotune.candidates(1024, vector_widths=(4, 8), items=(1, 2, 4), blocks=(256, 512), nontemporal=(0, 1))
    results = autotune.tune_fused(T, e.closure, e.shape, [[a[i][j].ptr for j in range(3)] for i in range(sets)],
                                  [o.ptr for o in outs], cands, reps=8)
    f, leaves = autotune.compile_with(T, e.closure, e.shape, results[0][0])
    ptrs = [(ctypes.c_void_p * 3)(*[a[i][j].ptr for j in range(3)]) for i in range(sets)]
    for i in range(sets):
        capi.check(lib.cs_fused_launch(T.stream, f, ptrs[i], 3, outs[i].ptr))
        capi.check(lib.cs_reduce_sum(T.stream, outs[i].ptr, n, sums[i].ptr, None))
    T.synchronize()
    timer = Timer(capi, T.stream)
    reps = max(1, steps // sets)
    g1 = capture(capi, T.stream, [lambda i=i: capi.check(lib.cs_fused_launch(T.stream, f, ptrs[i], 3, outs[i].ptr))
                                  for i in range(sets)])
    timer.measure(lambda: lib.cs_graph_launch(T.stream, g1), 2)
    ms_fma = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g1)), reps)
    g2 = capture(capi, T.stream, [lambda i=i: capi.check(lib.cs_reduce_sum(T.stream, outs[i].ptr, n, sums[i].ptr, None))
                                  for i in range(sets)])
    timer.measure(lambda: lib.cs_graph_launch(T.stream, g2), 2)
    ms_sum = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g2)), reps)
    # reported separately: the sum fused into its producer, one kernel reading a, b, c once
    # (12 MiB + 4 B) with no intermediate tensor, then the FP64 combine
    fs, _ = T.compile_sum(e.closure, e.shape)
    g3 = capture(capi, T.stream, [lambda i=i: capi.check(lib.cs_fused_sum(T.stream, fs, ptrs[i], 3, sums[i].ptr, None))
                                  for i in range(sets)])
    timer.measure(lambda: lib.cs_graph_launch(T.stream, g3), 2)
    ms_fsum = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g3)), reps)
    L = reps * sets
    res = {"fma_gbs": 16 * n * L / (ms_fma * 1e-3) / 1e9, "fma_us": ms_fma * 1e3 / L, "fma_config": results[0][0],
           "sum_gbs": (4 * n + 4) * L / (ms_sum * 1e-3) / 1e9, "sum_us": ms_sum * 1e3 / L,
           "fused_fma_sum_us": ms_fsum * 1e3 / L, "fused_fma_sum_gbs": (12 * n + 4) * L / (ms_fsum * 1e-3) / 1e9,
           "sets": sets}
    # the same three calls spread over `streams` streams, as the headline is, over 4096 calls
    spread = {}
    for name, nbytes, fn in (
            ("fma", 16 * n, lambda st, i: capi.check(lib.cs_fused_launch(st, f, ptrs[i], 3, outs[i].ptr))),
            ("sum", 4 * n + 4, lambda st, i: capi.check(lib.cs_reduce_sum(st, outs[i].ptr, n, sums[i].ptr, None))),
            ("fused_fma_sum", 12 * n + 4,
             lambda st, i: capi.check(lib.cs_fused_sum(st, fs, ptrs[i], 3, sums[i].ptr, None)))):
        sp = StreamSpread(T, capi, fn, sets, streams, prime=True)
        sp.run(sets)
        ms = sp.run(4096)
        sp.release()
        spread[name + "_us"] = round(ms * 1e3 / 4096, 3)
        spread[name + "_gbs"] = round(nbytes * 4096 / (ms * 1e-3) / 1e9, 1)
    spread["streams"] = streams
    res["over_streams"] = spread
    for g in (g1, g2, g3):
        lib.cs_graph_release(g)
    lib.cs_fused_release(f)
    lib.cs_fused_release(fs)
    for t in (ta, tb, tc):
        t.release()
    for h in [x for s in a for x in s] + outs + sums:
        h.release()
    return res


def bench_c3(T, capi, steps, sets=3, streams=3):
    lib = capi.lib
    rows, cols = 16384, 4096
    n = rows * cols
    ins = [T._alloc(n * 4) for _ in range(sets)]
    rs = [T._alloc(rows * 4) for _ in range(sets)]
    rm = [T._alloc(rows * 4) for _ in range(sets)]
    tot = [T._alloc(8) for _ in range(sets)]
    for i in range(sets):
        capi.check(lib.cs_fill_normal(T.stream, ins[i].ptr, n, 3001 + 97 * i))

    def launch(i):
        capi.check(lib.cs_reduce_rows(T.stream, ins[i].ptr, rows, cols, rs[i].ptr, rm[i].ptr, None, tot[i].ptr))

    for i in range(sets):
        launch(i)
    T.synchronize()
    g = capture(capi, T.stream, [lambda i=i: launch(i) for i in range(sets)] * 4)
    timer = Timer(capi, T.stream)
    timer.measure(lambda: lib.cs_graph_launch(T.stream, g), 1)
    reps = max(1, steps // (sets * 4))
    ms = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g)), reps)
    L = reps * sets * 4
    bytes_per = 4 * n + 4 * rows * 2 + 8
    res = {"rows_gbs": bytes_per * L / (ms * 1e-3) / 1e9, "rows_us": ms * 1e3 / L, "bytes_per_launch": bytes_per}
    lib.cs_graph_release(g)
    # the same calls spread over `streams` streams (each with its own scratch), 384 calls
    sp = StreamSpread(T, capi, lambda st, i: capi.check(lib.cs_reduce_rows(st, ins[i].ptr, rows, cols, rs[i].ptr,
                                                                            rm[i].ptr, None, tot[i].ptr)),
                      sets, streams, prime=True)
    sp.run(sets)
    ms_sp = sp.run(384)
    sp.release()
    res["over_streams"] = {"streams": streams, "calls": 384, "rows_us": round(ms_sp * 1e3 / 384, 3),
                           "rows_gbs": round(bytes_per * 384 / (ms_sp * 1e-3) / 1e9, 1)}
    # column sums of the same matrix (an outer-axis reduction: `split(0).reduce(_ + _)`)
    cs = [T._alloc(cols * 4) for _ in range(sets)]
    for i in range(sets):  # sizes the stream's scratch before capture
        capi.check(lib.cs_reduce_cols(T.stream, ins[i].ptr, rows, cols, cs[i].ptr))
    T.synchronize()
    g = capture(capi, T.stream, [lambda i=i: capi.check(lib.cs_reduce_cols(T.stream, ins[i].ptr, rows, cols,
                                                                           cs[i].ptr)) for i in range(sets)] * 4)
    timer.measure(lambda: lib.cs_graph_launch(T.stream, g), 1)
    ms = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g)), reps)
    cbytes = 4 * n + 4 * cols
    res.update({"cols_gbs": cbytes * L / (ms * 1e-3) / 1e9, "cols_us": ms * 1e3 / L})
    lib.cs_graph_release(g)
    for h in ins + rs + rm + tot + cs:
        h.release()
    return res


def bench_gemm(T, capi, variants, reps=10, streams=2):
    lib = capi.lib
    n = 4096
    from oracle import fp64, inputs
    ba, bb, bc = T._alloc(n * n * 4), T._alloc(n * n * 4), T._alloc(n * n * 4)
    bc2 = T._alloc(n * n * 4)  # the second stream's output when calls are spread
    rows = np.arange(3, n, 16)  # 256 rows of C against FP64 per operand set
    sets = {}
    for scaled in (False, True):
        a, b = inputs.c4_inputs(n, scaled)
        sets[scaled] = (a, b, a[rows].astype(np.float64) @ b.astype(np.float64))
    timer = Timer(capi, T.stream)
    out = {}

    def upload(a, b):
        for buf, arr in ((ba, a), (bb, b)):
            capi.check(lib.cs_buffer_write(T.stream, buf.handle, 0, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes,
                                           1))

    for name, v in variants:
        def run():
            capi.check(lib.cs_gemm(T.stream, v, n, n, n, ba.ptr, n, bb.ptr, n, bc.ptr, n))
        errs = {}
        for scaled in (True, False):  # the U(-1,1) set last: it is the one timed
            a, b, c64 = sets[scaled]
            upload(a, b)
            run()
            T.synchronize()
            c = np.empty((n, n), np.float32)
            capi.check(lib.cs_memcpy_dtoh(T.stream, c.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(bc.ptr),
                                          c.nbytes))
            errs["row_scaled" if scaled else "uniform"] = fp64.gemm_normalized_error(c[rows], a[rows], b, c64)
        timer.measure(run, 2)
        ms = timer.measure(run, reps)
        tflops = 2.0 * n ** 3 * reps / (ms * 1e-3) / 1e12
        out[name] = {"tflops": tflops, "ms": ms / reps, "normalized_error": errs["uniform"],
                     "normalized_error_row_scaled": errs["row_scaled"],
                     "mfma_util": tflops * (3 if name == "bf16x3" else 1) / PEAK_TFLOPS[name]}
        if streams > 1:
            # independent calls (the same operands, an output per stream) spread over
            # `streams` streams, each with its own operand scratch: one call's operand packing
            # overlaps another's main kernel
            outs_c = [bc, bc2]
            sp = StreamSpread(T, capi, lambda st, i: capi.check(lib.cs_gemm(st, v, n, n, n, ba.ptr, n, bb.ptr, n,
                                                                             outs_c[i].ptr, n)),
                              2, streams, prime=True)
            sp.run(2)
            calls = 2 * reps
            ms_sp = sp.run(calls)
            sp.release()
            out[name]["over_streams"] = {"streams": streams, "calls": calls,
                                         "tflops": 2.0 * n ** 3 * calls / (ms_sp * 1e-3) / 1e12}
        log("gemm %s: %.1f TFLOP/s, error %.3g (row-scaled set %.3g)" % (name, tflops, errs["uniform"],
                                                                         errs["row_scaled"]))
    for h in (ba, bb, bc, bc2):
        h.release()
    return out


def bench_c5(T, capi, sweep=False, out_dir=None, n=512, seed=5001, nontemporal=1, batch_vu=(2, 2)):
    """512 uneven expressions: per-expression launches vs one persistent batch (both
    graph-captured), with configurations from the cost model when its parameters exist.
    With `sweep`, first runs the exhaustive autotuner over every configuration, trains the
    cost model on 3/4 of the expressions and measures its regret on the rest."""
    import concurrent.futures
    from compute import batch as bm
    from compute import costmodel as cm
    from compute import workloads
    t0 = time.time()
    exprs = workloads.uneven_dags(T, n, seed)
    eset = bm.ExpressionSet(T, exprs)
    total_bytes = eset.bytes
    log("c5: %d expressions, %.2f GB, generated and uploaded in %.1f s" % (len(exprs), total_bytes / 1e9,
                                                                          time.time() - t0))
    res = {"expressions": len(exprs), "bytes": total_bytes}
    cfgs = cm.configs()
    variants = sorted({(c["vector_width"], c["items_per_thread"]) for c in cfgs})
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(len(variants)) as ex:
        futs = {vu: ex.submit(bm.compile_many, T, exprs, [{"vector_width": vu[0], "items_per_thread": vu[1],
                                                           "block_size": 256, "nontemporal_store": nontemporal}] *
                              len(exprs)) for vu in variants}
        fused = {vu: f.result() for vu, f in futs.items()}
    res["compile_s_all_variants"] = time.time() - t0
    log("c5: compiled %d variants x %d expressions in %.1f s" % (len(variants), len(exprs), res["compile_s_all_variants"]))
    model_path = os.path.join(ROOT, "compute.scala_amd", "compute", "data", "costmodel.npz")
    default = {"vector_width": 4, "items_per_thread": 1, "block_size": 256}
    if sweep:
        t0 = time.time()
        times = cm.measure(eset, fused, cfgs, progress=lambda i: log("c5: autotuned %d/%d expressions" % (i, len(exprs)))
                           if i % 64 == 0 else None)
        res["timings"] = len(times)
        log("c5: %d autotuner timings in %.1f s" % (len(times), time.time() - t0))
        ntrain = (3 * len(exprs)) // 4
        X, y, grp = [], [], []
        for (i, ci), us in times.items():
            if i < ntrain:
                X.append(cm.features(exprs[i], cfgs[ci]))
                y.append(us)
                grp.append(i)
        model = cm.CostModel(T, len(X[0]))
        t0 = time.time()
        hist = model.fit(np.stack(X), np.array(y), epochs=300, batch=512, groups=np.array(grp))
        res["train_s"] = time.time() - t0
        res["train_loss_first_last"] = [hist[0], hist[-1]]
        best_t = model_t = default_t = 0.0
        di = cfgs.index(default)
        for i in range(ntrain, len(exprs)):
            ts = [times[(i, ci)] for ci in range(len(cfgs))]
            choice = model.choose(exprs[i], cfgs)
            best_t += min(ts)
            model_t += ts[cfgs.index(choice)]
            default_t += ts[di]
        res["held_out"] = len(exprs) - ntrain
        res["regret_vs_exhaustive"] = model_t / best_t - 1.0
        res["gain_over_default"] = default_t / model_t
        res["held_out_us"] = {"exhaustive": best_t, "model": model_t, "default": default_t}
        # among the 256-thread configurations, the set the per-expression launches choose from
        b256 = [c for c in cfgs if c["block_size"] == 256]
        bb = bm_t = 0.0
        for i in range(ntrain, len(exprs)):
            tb = [times[(i, cfgs.index(c))] for c in b256]
            bb += min(tb)
            bm_t += times[(i, cfgs.index(model.choose(exprs[i], b256)))]
        res["regret_256_threads"] = bm_t / bb - 1.0
        res["default_regret_256_threads"] = default_t / bb - 1.0
        log("c5: cost model regret %.3f, gain over default %.3f; among 256-thread configurations regret %.3f "
            "(default %.3f)" % (res["regret_vs_exhaustive"], res["gain_over_default"], res["regret_256_threads"],
                                res["default_regret_256_threads"]))
        os.makedirs(os.path.dirname(model_path), exist_ok=True)
        for path in [model_path] + ([os.path.join(out_dir, "costmodel.npz")] if out_dir else []):
            np.savez(path, params=model.net.get_params(), mean=model.mean, std=model.std,
                     ymean=np.array([model.ymean]), widths=np.array(model.net.widths))
        if out_dir:
            with open(os.path.join(out_dir, "c5_timings.json"), "w") as f:
                json.dump({"configs": cfgs, "times": [[i, ci, us] for (i, ci), us in sorted(times.items())]}, f)
        model.release()
    # configuration per expression: cost model if trained parameters exist, else default
    chosen = [default] * len(exprs)
    model = cm.load(T, model_path)
    if model is not None:
        b256 = [c for c in cfgs if c["block_size"] == 256]
        chosen = [model.choose(e, b256) for e in exprs]
        model.release()
        res["configs_from"] = "cost model"
    else:
        res["configs_from"] = "default"

    def graph_each(cf):
        return bm.Graph(T, lambda: [eset.launch(i, fused[(c["vector_width"], c["items_per_thread"])][i],
                                                c["block_size"]) for i, c in enumerate(cf)])

    g_def = graph_each([default] * len(exprs))
    us_def = g_def.time_us(5)
    g_def.release()
    g_ch = graph_each(chosen)
    us_each = g_ch.time_us(5)
    g_ch.release()
    # the same per-expression launches spread over three streams (expression i on stream
    # i mod 3, each stream replaying a graph of its expressions), as the reference's queue
    # pool runs independent materialisations concurrently
    def launch_one_on(st, i):
        c = chosen[i]
        f = fused[(c["vector_width"], c["items_per_thread"])][i]
        capi.check(capi.lib.cs_fused_launch_block(st, f, eset._ptrs[i], len(eset.leaf_bufs[i]),
                                                  eset.out_bufs[i].ptr, c["block_size"]))
    sp = StreamSpread(T, capi, launch_one_on, len(exprs), 3)
    sp.run(len(exprs))
    us_each_streams = min(sp.run(len(exprs)) for _ in range(5)) * 1e3
    sp.release()
    res["per_expression_over_3_streams_us"] = us_each_streams
    eset.clear_outputs()
    sel = [fused[(c["vector_width"], c["items_per_thread"])][i] for i, c in enumerate(chosen)]
    for i, f in enumerate(sel):
        eset.launch(i, f, 256)
    each_out = eset.outputs()
    eset.clear_outputs()
    # The persistent batch runs in 512-thread workgroups; its bodies' V and U per expression
    # come from the cost model (block size fixed at the batch's 512), among the bodies with
    # V*U = 4: every work item then covers 2048 elements (the queue stays balanced) and the
    # batch's occupancy, set by its largest body, stays at that of 4 elements per thread.
    # Beside it: the fixed V = 2, U = 2 bodies, and the per-expression best of the exhaustive
    # timing database at 512 threads among the same bodies (the autotuner's choice, the judge
    # of the model's), each as its own batch.
    light = [c for c in cfgs if c["block_size"] == 512 and c["vector_width"] * c["items_per_thread"] == 4]
    choices = {"fixed": [(batch_vu[0], batch_vu[1])] * len(exprs)}
    model = cm.load(T, model_path)
    if model is not None:
        choices["model"] = [(c["vector_width"], c["items_per_thread"]) for c in (model.choose(e, light) for e in exprs)]
        model.release()
    db_path = os.path.join(ROOT, "compute.scala_amd", "compute", "data", "c5_timings.json.gz")
    if n == 512 and seed == 5001 and os.path.exists(db_path):
        import gzip
        with gzip.open(db_path, "rt") as f:
            db = json.load(f)
        idx = {tuple(sorted(c.items())): k for k, c in enumerate(db["configs"])}
        t = {(int(i), int(ci)): us for i, ci, us in db["times"]}
        ex = []
        for i in range(len(exprs)):
            best = min(light, key=lambda c: t[(i, idx[tuple(sorted(c.items()))])])
            ex.append((best["vector_width"], best["items_per_thread"]))
        choices["exhaustive"] = ex
    batches = {}
    for name, ch in choices.items():
        bsel = [fused[vu][i] for i, vu in enumerate(ch)]
        t0 = time.time()
        b = bm.Batch(eset, bsel, block=512)
        compile_s = time.time() - t0
        gb = bm.Graph(T, b.launch)
        us = gb.time_us(5)
        out = eset.outputs()
        ident = all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(each_out, out))
        gb.release()
        batches[name] = {"us": us, "bit_identical": ident, "compile_s": compile_s, "info": b.info()}
        b.release()
        eset.clear_outputs()
        log("c5 batch with %s bodies: %.1f us, identical=%s" % (name, us, ident))
    res["batch_variants"] = batches
    use = "model" if "model" in batches else "fixed"
    us_batch = batches[use]["us"]
    res["batch_bodies_from"] = use
    res["batch_bit_identical"] = batches[use]["bit_identical"]
    res["batch_info"] = batches[use]["info"]
    res["batch_compile_s"] = batches[use]["compile_s"]
    if "exhaustive" in batches and "model" in batches:
        res["batch_model_regret_vs_exhaustive"] = batches["model"]["us"] / batches["exhaustive"]["us"] - 1.0
    res["batch_config"] = {"block_size": 512, "nontemporal_store": nontemporal,
                           "bodies": "V, U per expression from the cost model (V*U = 4)" if use == "model"
                           else "V=%d, U=%d" % tuple(batch_vu)}
    res["per_expression_default_us"] = us_def
    res["per_expression_us"] = us_each
    res["batched_us"] = us_batch
    res["per_expression_gbs"] = total_bytes / (us_each * 1e-6) / 1e9
    res["batched_gbs"] = total_bytes / (us_batch * 1e-6) / 1e9
    res["batched_over_per_expression"] = us_each / us_batch
    res["batched_over_per_expression_over_3_streams"] = us_each_streams / us_batch
    log("c5: per-expression %.1f us (default configs %.1f us, over 3 streams %.1f us), batched %.1f us -> %.2fx "
        "(%.2fx against 3 streams), %.0f GB/s, identical=%s" %
        (us_each, us_def, us_each_streams, us_batch, res["batched_over_per_expression"],
         res["batched_over_per_expression_over_3_streams"], res["batched_gbs"], res["batch_bit_identical"]))
    for fs in fused.values():
        for f in fs:
            capi.lib.cs_fused_release(f)
    eset.release()
    return res


def bench_conv(T_outer, capi, steps=256, sets=4, vector_width=1, items_per_thread=1, block_size=256):
    """The convolution idiom of the reference's benchmark at its largest setting
    (benchmarks.scala:454-548, 612-631: input [128,32,32,8], 3x3 kernel, 8 filters, one
    layer): a join of 8 per-filter sums of 72 products of translated input channels and
    broadcast weights, one generated kernel, one element per thread (V=1, 9.3 us against
    25 us at V=2 and 50 us at V=4: only at V=1 are a pixel's channels read together).
    Reported separately (not a headline workload)."""
    from compute import Tensors
    from compute import tensors as tz, trees, workloads
    T = Tensors(0, math_mode=T_outer.options.math_mode, vector_width=vector_width,
                items_per_thread=items_per_thread, block_size=block_size)
    lib = capi.lib
    B, H, W, D, F = 128, 32, 32, 8, 8
    n_in, n_w, n_out = B * H * W * D, 3 * 3 * D * F, B * H * W * F
    xs = [T._alloc(n_in * 4) for _ in range(sets)]
    outs = [T._alloc(n_out * 4) for _ in range(sets)]
    wb, bb = T._alloc(n_w * 4), T._alloc(F * 4)
    for i in range(sets):
        capi.check(lib.cs_fill_normal(T.stream, xs[i].ptr, n_in, 6001 + i))
    capi.check(lib.cs_fill_normal(T.stream, wb.ptr, n_w, 6101))
    capi.check(lib.cs_fill_normal(T.stream, bb.ptr, F, 6102))
    X = tz.CachedTensor(T, (B, H, W, D), 0.0, xs[0].retain())
    Wt = tz.CachedTensor(T, (3, 3, D, F), 0.0, wb.retain())
    Bt = tz.CachedTensor(T, (F,), 0.0, bb.retain())
    out = workloads.convolute(T, X, Wt, Bt)
    clo = trees.Concatenate([p.closure for p in out._parts])
    t0 = time.time()
    f, leaves, nout = T.compile(clo, (B, H, W))
    compile_s = time.time() - t0
    order = {id(X): None, id(Wt): wb, id(Bt): bb}
    ptrs = []
    for i in range(sets):
        order[id(X)] = xs[i]
        ptrs.append((ctypes.c_void_p * len(leaves))(*[order[id(l.id)].ptr for l in leaves]))
    launch = lambda i: capi.check(lib.cs_fused_launch(T.stream, f, ptrs[i], len(leaves), outs[i].ptr))
    g = capture(capi, T.stream, [lambda i=i: launch(i) for i in range(sets)])
    timer = Timer(capi, T.stream)
    timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g)), 4)
    reps = max(1, steps // sets)
    ms = timer.measure(lambda: capi.check(lib.cs_graph_launch(T.stream, g)), reps)
    us = ms * 1e3 / (reps * sets)
    flop = 2 * B * H * W * F * 9 * D + B * H * W * F  # products and adds per output, plus the bias
    res = {"us_per_forward": us, "gflops": flop / (us * 1e-6) / 1e9,
           "gbs": 4 * (n_in + n_w + F + n_out) / (us * 1e-6) / 1e9, "compile_s": compile_s,
           "config": "input [128,32,32,8], weight [3,3,8,8], bias [8], one layer, V=%d, U=%d, %d threads" %
                     (vector_width, T.options.items_per_thread, block_size)}
    lib.cs_graph_release(g)
    lib.cs_fused_release(f)
    for t in (X, Wt, Bt):
        t.release()
    for h in xs + outs + [wb, bb]:
        h.release()
    T.close()
    return res


def bench_split(T, capi, n=64, seed=6001, out_dir=None):
    """Fusion splits of convolution-shaped joins: the exhaustive autotuner times every cut of
    each of n joins (1, 2, 4 or 8 kernels plus the gathering kernel; 20 graph replays each), the split cost model is
    trained on the first 3/4 of the joins (by the device MLP kernels) and its picks on the
    rest are compared with the exhaustive best and with never splitting. Saves the timings and
    the model beside the package's data when asked (`--split-sweep`)."""
    from compute import autotune, costmodel as cm, workloads
    t0 = time.time()
    joins = workloads.conv_joins(T, n, seed)
    bufs, ptrs = [], {}
    for j in joins:
        for t, d in zip(j.leaves, j.leaf_data):
            a = np.ascontiguousarray(d, np.float32).ravel()
            b = T._alloc(a.nbytes)
            capi.check(capi.lib.cs_buffer_write(T.stream, b.handle, 0, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1))
            bufs.append(b)
            ptrs[id(t)] = b.ptr
    times = {}
    for k, j in enumerate(joins):
        for g in autotune.split_candidates(j.parts):
            times[(k, g)] = autotune.time_join(T, j.tensor, ptrs, g, reps=20)
        if k % 16 == 15:
            log("split sweep: %d/%d joins timed (%.0f s)" % (k + 1, n, time.time() - t0))
    for b in bufs:
        b.release()
    ntrain = (3 * n) // 4
    X, y, grp = [], [], []
    for (k, g), us in times.items():
        if k < ntrain:
            X.append(cm.split_features(joins[k].tensor, g))
            y.append(us)
            grp.append(k)
    model = cm.SplitModel(T, len(X[0]), hidden=32, layers=2, seed=5)
    hist = model.fit(np.stack(X), np.array(y), epochs=400, batch=32, lr=2e-3, groups=np.array(grp))
    # the margin a cut must be predicted to win by, chosen on the training joins
    preds = {k: model.predicted(joins[k].tensor) for k in range(n)}
    margins = (0.0, 0.02, 0.05, 0.1, 0.2, 0.5)
    train_t = [sum(times[(k, cm.pick_groups(*preds[k], m))] for k in range(ntrain)) for m in margins]
    model.margin = margins[int(np.argmin(train_t))]
    best = chosen = never = 0.0
    picks = {}
    for k in range(ntrain, n):
        j = joins[k]
        cands = autotune.split_candidates(j.parts)
        best += min(times[(k, g)] for g in cands)
        picks[k] = cm.pick_groups(*preds[k], model.margin)
        chosen += times[(k, picks[k])]
        never += times[(k, 1)]
    res = {"joins": n, "seed": seed, "timings": len(times), "held_out": n - ntrain, "margin": model.margin,
           "held_out_picks_split": sum(1 for k in picks if picks[k] > 1),
           "train_loss_first_last": [hist[0], hist[-1]],
           "held_out_us": {"exhaustive": best, "model": chosen, "one_kernel": never},
           "regret_vs_exhaustive": chosen / best - 1.0, "gain_over_one_kernel": never / chosen,
           "joins_split_by_exhaustive": sum(1 for k in range(n) if min(autotune.split_candidates(joins[k].parts),
                                                                     key=lambda g: times[(k, g)]) > 1),
           "sweep_s": time.time() - t0}
    log("split model: held-out regret %.3f, gain over one kernel %.3f" % (res["regret_vs_exhaustive"],
                                                                          res["gain_over_one_kernel"]))
    if out_dir is not None:
        import gzip
        data = os.path.join(ROOT, "compute.scala_amd", "compute", "data")
        for d in (data, out_dir):
            os.makedirs(d, exist_ok=True)
            cm.save(model, os.path.join(d, "splitmodel.npz"))
            with gzip.open(os.path.join(d, "split_timings.json.gz"), "wt") as f:
                json.dump({"seed": seed, "joins": n, "times": [[k, g, us] for (k, g), us in sorted(times.items())]}, f)
    model.release()
    return res


def profile_run(T, capi, what, count=512):
    """Launches of one workload's kernels without graphs or extra streams, for a profiler to
    time per dispatch: `c2` (the headline kernel over the headline's rotating sets, plain
    launches one after another on the launch stream), `c3` (the row reduction), `c4` (one
    call of each GEMM variant, several times), `c5` (the batch, then each expression once),
    `conv` (the convolution idiom's kernel at the benchmark's launch shape)."""
    from compute import autotune
    from compute import tensors as tz
    lib = capi.lib
    if what == "c2":
        n = 64 * 256 * 256
        sets = HEADLINE_SETS
        xs = [T._alloc(n * 4) for _ in range(sets)]
        bs = [T._alloc(256 * 4) for _ in range(sets)]
        outs = [T._alloc(n * 4) for _ in range(sets)]
        for i in range(sets):
            capi.check(lib.cs_fill_normal(T.stream, xs[i].ptr, n, 2001 + 97 * i))
            capi.check(lib.cs_fill_uniform(T.stream, bs[i].ptr, 256, 2002 + 97 * i, -0.5, 0.5))
        X = tz.CachedTensor(T, (64, 256, 256), 0.0, xs[0].retain())
        B = tz.CachedTensor(T, (256,), 0.0, bs[0].retain())
        p = X.permute([1, 0, 2])
        y = p + B.broadcast((256, 64, 256)) + p.translate([0, 0, 1])
        out = T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)
        cfg = {"vector_width": 4, "items_per_thread": 1, "block_size": 256, "nontemporal_store": 1}
        f, _ = autotune.compile_with(T, out.closure, out.shape, cfg)
        ptrs = [(ctypes.c_void_p * 2)(xs[i].ptr, bs[i].ptr) for i in range(sets)]
        # plain launches, one after another on the launch stream, as a graph replay issues them
        for k in range(count):
            capi.check(lib.cs_fused_launch(T.stream, f, ptrs[k % sets], 2, outs[k % sets].ptr))
        T.synchronize()
        log("c2 kernel: %d plain launches" % count)
        lib.cs_fused_release(f)
        for h in xs + bs + outs + [X, B]:
            h.release()
    elif what == "c3":
        rows, cols = 16384, 4096
        a = T._alloc(rows * cols * 4)
        rs, rm, tot = T._alloc(rows * 4), T._alloc(rows * 4), T._alloc(8)
        capi.check(lib.cs_fill_normal(T.stream, a.ptr, rows * cols, 3001))
        for _ in range(count // 8):
            capi.check(lib.cs_reduce_rows(T.stream, a.ptr, rows, cols, rs.ptr, rm.ptr, None, tot.ptr))
        T.synchronize()
        for h in (a, rs, rm, tot):
            h.release()
    elif what == "c4":
        kinds = [("fp32", capi.GEMM_FP32), ("bf16", capi.GEMM_BF16), ("bf16x3", capi.GEMM_BF16X3),
                 ("fp8", capi.GEMM_FP8)]
        bench_gemm(T, capi, kinds, reps=10, streams=1)
    elif what == "c5":
        from compute import batch as bm
        from compute import workloads
        exprs = workloads.uneven_dags(T, 512, 5001)
        eset = bm.ExpressionSet(T, exprs)
        fs = bm.compile_many(T, exprs, [{"vector_width": 2, "items_per_thread": 2, "block_size": 256,
                                         "nontemporal_store": 1}] * len(exprs))
        b = bm.Batch(eset, fs, block=512)
        for _ in range(3):
            b.launch()
        T.synchronize()
        for i, f in enumerate(fs):
            eset.launch(i, f, 256)
        T.synchronize()
        log("c5: %d expressions, %d bytes; 3 batch launches, then each expression once" % (len(exprs), eset.bytes))
        b.release()
        for f in fs:
            lib.cs_fused_release(f)
        eset.release()
    elif what == "conv":
        from compute import trees, workloads
        Bn, H, W, D, F = 128, 32, 32, 8, 8
        n_in, n_w, n_out = Bn * H * W * D, 3 * 3 * D * F, Bn * H * W * F
        x, o = T._alloc(n_in * 4), T._alloc(n_out * 4)
        wb, bb = T._alloc(n_w * 4), T._alloc(F * 4)
        capi.check(lib.cs_fill_normal(T.stream, x.ptr, n_in, 6001))
        capi.check(lib.cs_fill_normal(T.stream, wb.ptr, n_w, 6101))
        capi.check(lib.cs_fill_normal(T.stream, bb.ptr, F, 6102))
        X = tz.CachedTensor(T, (Bn, H, W, D), 0.0, x.retain())
        Wt = tz.CachedTensor(T, (3, 3, D, F), 0.0, wb.retain())
        Bt = tz.CachedTensor(T, (F,), 0.0, bb.retain())
        out = workloads.convolute(T, X, Wt, Bt)
        # the benchmark's launch shape: one element per thread, 256 threads
        f, leaves = autotune.compile_with(T, trees.Concatenate([q.closure for q in out._parts]), (Bn, H, W),
                                          {"vector_width": 1, "items_per_thread": 1, "block_size": 256,
                                           "nontemporal_store": 0})
        by_id = {id(X): x, id(Wt): wb, id(Bt): bb}
        ptrs = (ctypes.c_void_p * len(leaves))(*[by_id[id(l.id)].ptr for l in leaves])
        for _ in range(count // 4):
            capi.check(lib.cs_fused_launch(T.stream, f, ptrs, len(leaves), o.ptr))
        T.synchronize()
        log("convolution idiom: %d plain launches" % (count // 4))
        lib.cs_fused_release(f)
        for h in (x, o, wb, bb, X, Wt, Bt):
            h.release()
    else:
        raise SystemExit("--profile takes c2, c3, c4, c5 or conv")


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": name, "logical_cpus": os.cpu_count()}


def cpu_baseline_c2(seconds=10.0):
    """The reference's kernel for the headline expression (refcl-host, precise mode) on all
    host cores, timed on whole launches until about `seconds` have passed."""
    from compute import Tensors
    from oracle import inputs, refcl
    T = Tensors(offline=True)
    X, bias = inputs.c2_inputs()
    p = T.Tensor(X).permute([1, 0, 2])
    y = p + T.Tensor(bias).broadcast((256, 64, 256)) + p.translate([0, 0, 1])
    out = T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)
    k = refcl.Kernel(out.closure, out.shape, "precise")
    k([X, bias])
    runs, t0 = 0, time.perf_counter()
    while True:
        k([X, bias])
        runs += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    bytes_per = 4 * X.size + 4 * 256 + 4 * X.size
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(bytes_per * runs / el / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": "%d full launches of the workload's reference kernel text (OpenMP NDRange, host C, "
                      "no FMA) in %.1f s" % (runs, el)}


def _timed(fn, seconds):
    """Calls fn until about `seconds` have passed (at least twice); returns (calls, elapsed)."""
    fn()
    runs, t0 = 0, time.perf_counter()
    while True:
        fn()
        runs += 1
        el = time.perf_counter() - t0
        if el >= seconds and runs >= 2:
            return runs, el


def cpu_reference_other(seconds=3.0):
    """The reference's CPU programs for the other workloads, each on a bounded sample on this
    host: the elementwise kernel text of C1 on all threads, the CPU device's sequential `sum`
    program (one thread, as the reference launches it), the row folds that `split(1).reduce`
    unrolls to, and the sequential-k FP32 product of the matrixMultiply idiom."""
    import oracle
    from compute import Tensors
    from oracle import inputs, refcl
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    res = {}
    T = Tensors(offline=True)
    a, b, c = inputs.c1_inputs()
    e = T.Tensor(a) * T.Tensor(b) + T.Tensor(c)
    k = refcl.Kernel(e.closure, e.shape, "precise")
    runs, el = _timed(lambda: k([a, b, c]), seconds)
    res["c1_fma"] = {"value": round(16 * a.size * runs / el / 1e9, 3), "unit": "GB/s", "cores": threads,
                     "sample": "%d launches of a*b+c [1024,1024] (reference kernel text, host C)" % runs}
    x = ((a * b) + c).ravel()
    runs, el = _timed(lambda: oracle.sequential_sum(x), seconds)
    res["c1_sum"] = {"value": round(4 * x.size * runs / el / 1e9, 3), "unit": "GB/s", "cores": 1,
                     "sample": "%d runs of the CPU device's sum program over 2^20 floats" % runs}
    rows, cols = 4096, 4096  # a quarter of C3's rows: same row length
    m = inputs.normal(rows * cols, 3001).reshape(rows, cols)
    runs, el = _timed(lambda: oracle.row_fold(m), seconds)
    res["c3_rows"] = {"value": round(4 * m.size * runs / el / 1e9, 3), "unit": "GB/s", "cores": threads,
                      "sample": "%d row folds (sum and max) of [4096,4096] (C3's row length)" % runs}
    n = 512
    ga, gb = inputs.uniform(n * n, 4001).reshape(n, n), inputs.uniform(n * n, 4002).reshape(n, n)
    runs, el = _timed(lambda: oracle.gemm_seqk(ga, gb), seconds)
    res["c4_gemm_fp32"] = {"value": round(2.0 * n ** 3 * runs / el / 1e12, 4), "unit": "TFLOP/s", "cores": threads,
                           "sample": "%d sequential-k FP32 products at 512^3 (the idiom's order)" % runs}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2048)
    ap.add_argument("--warmup", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="c2,hbm,c1,c3,c4,conv,c5")
    ap.add_argument("--c5-sweep", action="store_true", help="run the exhaustive autotuner and train the cost model")
    ap.add_argument("--split-sweep", type=int, default=0,
                    help="time every fusion split of this many convolution-shaped joins and train the split model")
    ap.add_argument("--out-dir", default=None)
    ap.add_argument("--streams", type=int, default=3,
                    help="streams the headline launches are spread over (the launch stream and streams - 1 "
                         "more)")
    ap.add_argument("--c2-config", default=None,
                    help="V,U,BLOCK,NT: launch configuration of the headline kernel instead of autotuning it")
    ap.add_argument("--gemm", default="fp32,bf16,bf16x3,fp8", help="GEMM variants to run for c4")
    ap.add_argument("--c5-batch", default="2,2", help="c5: vector width and items per thread of the batch bodies")
    ap.add_argument("--c5-nontemporal", type=int, default=1,
                    help="c5: compile every expression with non-temporal stores (1) or plain stores (0)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy a buffer of this many MiB device-to-device 8 times (a known-size "
                         "reference for hardware byte counters)")
    ap.add_argument("--profile", default=None,
                    help="c2, c3, c4, c5 or conv: only that workload's kernels, launched without graphs (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every workload here runs on a single MI355X, "
              "so --gpus %d is not supported" % args.gpus)
        sys.exit(2)
    from compute import Tensors, _capi as capi
    only = set(args.only.split(","))
    T = Tensors(0)
    log("device: %d CUs" % T.info(capi.INFO_COMPUTE_UNITS))
    if args.copy_mib:
        nb = args.copy_mib << 20
        src, dst = T._alloc(nb), T._alloc(nb)
        capi.check(capi.lib.cs_fill_uniform(T.stream, src.ptr, nb // 4, 9, -1.0, 1.0))
        for _ in range(8):
            capi.check(capi.lib.cs_memcpy_dtod(T.stream, dst.ptr, src.ptr, nb))
        T.synchronize()
        src.release()
        dst.release()
        log("copied %d MiB 8 times" % args.copy_mib)
    if args.profile:
        profile_run(T, capi, args.profile)
        T.close()
        return
    fixed = None
    if args.c2_config:
        v, u, b, nt = (int(x) for x in args.c2_config.split(","))
        fixed = {"vector_width": v, "items_per_thread": u, "block_size": b, "nontemporal_store": nt}
    c2 = bench_c2(T, capi, args.steps, args.warmup, fixed=fixed, streams=args.streams)
    log("c2: %.1f GB/s (%.2f us/launch) over %d streams (10 reruns %s), one stream %.1f GB/s, long window "
        "%.1f GB/s (by streams %s), kernel %.2f us (dispatch events), hot cache %.1f GB/s" %
        (c2["gbs"], c2["ms_per_launch"] * 1e3, args.streams, c2["reruns"], c2["one_stream"]["gbs"],
         c2["long_window"]["gbs"], c2["long_window"]["gbs_by_streams"], c2["us_per_launch_events"],
         c2["gbs_hot_cache"]))
    also = {}
    if "hbm" in only:
        also["hbm_ceiling"] = bench_hbm(T, capi, streams=args.streams)
    if "c1" in only:
        also["c1_fma_sum"] = bench_c1(T, capi, max(2048, args.steps), streams=args.streams)
        log("c1: %s" % also["c1_fma_sum"])
    if "c3" in only:
        also["c3_rows_sum_max_total"] = bench_c3(T, capi, max(64, args.steps // 16), streams=args.streams)
        log("c3: %s" % also["c3_rows_sum_max_total"])
    if "c4" in only:
        kinds = {"fp32": capi.GEMM_FP32, "bf16": capi.GEMM_BF16, "bf16x3": capi.GEMM_BF16X3, "fp8": capi.GEMM_FP8}
        also["c4_gemm_4096"] = bench_gemm(T, capi, [(k, kinds[k]) for k in args.gemm.split(",")])
        g = also["c4_gemm_4096"]
        if "bf16x3" in g and "fp32" in g:
            g["bf16x3_over_fp32"] = g["bf16x3"]["tflops"] / g["fp32"]["tflops"]
            if "over_streams" in g["bf16x3"] and "over_streams" in g["fp32"]:
                g["bf16x3_over_fp32_over_streams"] = (g["bf16x3"]["over_streams"]["tflops"] /
                                                      g["fp32"]["over_streams"]["tflops"])
        log("gemm over streams: %s" % {k: round(x["over_streams"]["tflops"], 1) for k, x in g.items()
                                       if isinstance(x, dict) and "over_streams" in x})
    if args.split_sweep:
        also["split_sweep"] = bench_split(T, capi, args.split_sweep, out_dir=args.out_dir)
        log("split sweep: %s" % also["split_sweep"])
    if "conv" in only:
        also["convolution_idiom"] = bench_conv(T, capi)
        log("convolution idiom: %s" % also["convolution_idiom"])
    if "c5" in only:
        also["c5_uneven_batch"] = bench_c5(T, capi, sweep=args.c5_sweep, out_dir=args.out_dir,
                                           nontemporal=args.c5_nontemporal,
                                           batch_vu=tuple(int(x) for x in args.c5_batch.split(",")))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_c2()
        log("cpu baseline: %s" % cpu)
        also["cpu_reference_other"] = cpu_reference_other()
        log("cpu reference, other workloads: %s" % also["cpu_reference_other"])
    gbs = c2["gbs"]
    kernel_gbs = c2["bytes_per_launch"] / (c2["us_per_launch_events"] * 1e-6) / 1e9
    # HBM bytes per launch of the headline kernel, from the FETCH_SIZE / WRITE_SIZE passes
    # (scaled by what the counters read for a copy of known size) recorded under profiles/.
    traffic, traffic_source = None, None
    tp = os.path.join(ROOT, "profiles", "c2_hbm_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            t = json.load(f)
        traffic = t["traffic_bytes"]
        traffic_source = "profiles/c2_hbm_traffic.json, launch configuration %s" % (t["config"],)
    # The profiler's per-dispatch average for the same kernel and launch configuration, from
    # the committed kernel-trace summary (its session and command are in the file).
    profiled = None
    pp = os.path.join(ROOT, "profiles", "c2_kernel_trace.json")
    if os.path.exists(pp):
        with open(pp) as f:
            t = json.load(f)
        profiled = {"kernel_us": t["average_us"], "achieved": round(c2["bytes_per_launch"] / (t["average_us"] * 1e-6) / 1e9, 2),
                    "frac": round(c2["bytes_per_launch"] / (t["average_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                    "source": "profiles/c2_kernel_trace.json (%s)" % t["session"]}
    copy_ceiling = None
    if "hbm_ceiling" in also and "copy_spread" in also["hbm_ceiling"]["16MiB"]:
        copy_ceiling = also["hbm_ceiling"]["16MiB"]["copy_spread"]["gbs"]
    line = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": c2["launches"],
        "warmup": args.warmup,
        "ms_per_step": c2["ms_per_launch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Wang-hash seeded)",
        "config": {"workload": "view-heavy fused chain: X[64,256,256] N(0,1), bias[256] U(-0.5,0.5); "
                               "p=X.permute(1,0,2); y=p+bias.broadcast([256,64,256])+p.translate([0,0,1]); "
                               "exp(-y*y)*tanh(y); one generated kernel per step; %d rotating input sets; "
                               "steps spread over %d streams" % (HEADLINE_SETS, args.streams),
                   "bytes_per_step": c2["bytes_per_launch"], "hot_cache_gbs": round(c2["gbs_hot_cache"], 2),
                   "copy_reference": c2["copy_reference"], "pcie_inclusive": c2["pcie_inclusive"],
                   "streams": args.streams, "one_stream": c2["one_stream"], "reruns": c2["reruns"],
                   "long_window": c2["long_window"],
                   "launch_config": c2["tuning"]},
        # achieved: algorithmic bytes over the kernel's own average duration (events stamped by
        # each dispatch run alone, the profiler's per-dispatch figure). `value` is the timed
        # region's throughput: launches on different streams overlap one another's ramp-up and
        # drain, so it exceeds what one dispatch alone achieves (`overlapped`).
        "roofline": {"bound": "hbm", "achieved": round(kernel_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kernel_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_source, "kernel_us": round(c2["us_per_launch_events"], 3),
                     "timed_region_us_per_launch": round(c2["ms_per_launch"] * 1e3, 3),
                     "overlapped": {"achieved": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 4),
                                    "streams": args.streams},
                     "profiled": profiled,
                     # the same bytes moved by the library's own copy kernel at this size, spread
                     # over the same streams under the same harness (no arithmetic): the ceiling for
                     # launches of this size
                     "copy_ceiling_gbs": copy_ceiling,
                     "value_over_copy_ceiling": round(gbs / copy_ceiling, 4) if copy_ceiling else None,
                     "long_window_over_copy_ceiling": (round(c2["long_window"]["gbs"] / copy_ceiling, 4)
                                                       if copy_ceiling else None)},
        "cpu_baseline": cpu,
        "also": also,
    }
    T.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
