"""The runtime half of the C ABI: arena determinism, staging, programs, launches,
events, graphs and error mapping."""
import ctypes

import numpy as np
import pytest

from compute import _capi as capi
from gpu_util import alloc, download, upload
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu
lib = capi.lib


def test_device_info(gpu_tensors):
    T = gpu_tensors
    assert T.info(capi.INFO_DEVICE_TYPE) == 4
    assert T.info(capi.INFO_WAVEFRONT_SIZE) == 64
    assert T.info(capi.INFO_COMPUTE_UNITS) >= 1
    name = ctypes.c_char_p()
    capi.check(lib.cs_device_name(T.device, ctypes.byref(name)))
    assert name.value.startswith(b"gfx9")


def test_arena_is_deterministic(gpu_tensors):
    T = gpu_tensors

    def sequence():
        bufs = [alloc(T, n) for n in (1000, 1 << 20, 7, 1 << 16)]
        addrs = [b.ptr for b in bufs]
        bufs[1].release()
        c = alloc(T, 5000)
        addrs.append(c.ptr)
        for b in (bufs[0], bufs[2], bufs[3], c):
            b.release()
        return addrs

    T.synchronize()
    a1 = sequence()
    a2 = sequence()
    assert a1 == a2
    assert all(p % 256 == 0 for p in a1)


def test_materialisations_do_not_wait_for_the_device(gpu_tensors):
    """A chain of 100 facade materialisations (x <- x * 0.5 + 1 over 2^28 floats, each one a
    new buffer while the previous one is released) is queued without the host waiting for
    the device: the loop returns while the device still has part of the chain to run (its
    last kernel has not completed), the arena hands back the same two addresses in turn, and
    the result is the oracle's (checked on every 4099th element: the elements are
    independent)."""
    T = gpu_tensors
    n = 1 << 28
    x0 = inputs.uniform(n, 9001)
    t = T.Tensor(x0).doCache()
    T.synchronize()
    half, one = T.Tensor.fill(0.5, (n,)), T.Tensor.fill(1.0, (n,))
    import time
    t0 = time.perf_counter()
    addrs = []
    for _ in range(100):
        nxt = (t * half + one).doCache()
        addrs.append(nxt.buffer.ptr)
        t.release()
        t = nxt
    host_s = time.perf_counter() - t0
    ev = ctypes.c_void_p()
    capi.check(lib.cs_event_record(T.stream, ctypes.byref(ev)))
    done = ctypes.c_int()
    capi.check(lib.cs_event_query(ev, ctypes.byref(done)))
    t1 = time.perf_counter()
    capi.check(lib.cs_event_wait(ev))
    drain_s = time.perf_counter() - t1
    lib.cs_event_release(ev)
    print("100 materialisations queued in %.2f ms; the device finished %.2f ms later" % (host_s * 1e3, drain_s * 1e3))
    assert done.value == 0 and drain_s > 0.0
    # two buffers alternate: the same two addresses, in turn
    assert len(set(addrs)) == 2 and all(addrs[i] == addrs[i % 2] for i in range(100))
    want = x0[::4099].copy()
    for _ in range(100):
        want = want * np.float32(0.5) + np.float32(1.0)
    assert np.array_equal(t.flatArray()[::4099].view(np.uint32), want.view(np.uint32))
    t.release()


def test_staging_round_trip_large(gpu_tensors):
    T = gpu_tensors
    x = np.random.default_rng(0).standard_normal(40 << 20 >> 2).astype(np.float32)  # crosses staging chunks
    b = upload(T, x)
    assert np.array_equal(download(T, b, x.size), x)
    y = np.empty(1000, np.float32)
    capi.check(lib.cs_buffer_read(T.stream, b.handle, 4 * 123, y.ctypes.data_as(ctypes.c_void_p), 4000))
    assert np.array_equal(y, x[123:1123])
    with pytest.raises(capi.InvalidValue):
        capi.check(lib.cs_buffer_read(T.stream, b.handle, b.nbytes - 4, y.ctypes.data_as(ctypes.c_void_p), 8))
    b.release()


def test_non_blocking_write_frees_the_host_memory_at_return(gpu_tensors):
    """cs_buffer_write with blocking = 0 copies the host bytes into pinned staging before it
    returns, so the caller may overwrite its array at once: the device keeps what was passed.
    Sizes below and above the size at which the host copy is split over two threads."""
    T = gpu_tensors
    for nbytes in (100 << 10, (9 << 20) + 4):
        x = np.frombuffer(np.random.default_rng(nbytes).bytes(nbytes), np.uint8).copy()
        want = x.copy()
        b = alloc(T, nbytes // 4 + 1)
        capi.check(lib.cs_buffer_write(T.stream, b.handle, 0, x.ctypes.data_as(ctypes.c_void_p), nbytes, 0))
        x[:] = 0xA5  # overwritten right after the call returned
        y = np.zeros(nbytes, np.uint8)
        capi.check(lib.cs_buffer_read(T.stream, b.handle, 0, y.ctypes.data_as(ctypes.c_void_p), nbytes))
        assert np.array_equal(y, want)
        b.release()


def test_staging_wraps_its_slots(gpu_tensors):
    """100 MiB and 12,345 bytes: more chunks than pinned slots in both directions, and a short
    last chunk."""
    T = gpu_tensors
    nbytes = (100 << 20) + 12345
    x = np.frombuffer(np.random.default_rng(1).bytes(nbytes), np.uint8)
    b = alloc(T, nbytes // 4 + 1)
    capi.check(lib.cs_buffer_write(T.stream, b.handle, 0, x.ctypes.data_as(ctypes.c_void_p), nbytes, 0))
    y = np.zeros(nbytes, np.uint8)
    capi.check(lib.cs_buffer_read(T.stream, b.handle, 0, y.ctypes.data_as(ctypes.c_void_p), nbytes))
    assert np.array_equal(x, y)
    b.release()


SRC = b"""
extern "C" __global__ void axpy(float a, const float* x, float* y, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + y[i];
}
"""


def test_program_build_and_launch(gpu_tensors):
    T = gpu_tensors
    prog, log = ctypes.c_void_p(), ctypes.c_void_p()
    capi.check(lib.cs_program_build(T.device, SRC, b"-ffp-contract=off", ctypes.byref(prog), ctypes.byref(log)))
    capi.take_string(log.value)
    k = ctypes.c_void_p()
    capi.check(lib.cs_program_get_kernel(prog, b"axpy", ctypes.byref(k)))
    with pytest.raises(capi.InvalidKernelName):
        capi.check(lib.cs_program_get_kernel(prog, b"nope", ctypes.byref(ctypes.c_void_p())))
    mb = ctypes.c_int(0)
    capi.check(lib.cs_kernel_max_block(k, ctypes.byref(mb)))
    assert 256 <= mb.value <= 1024  # kernel.workGroupSize(deviceId)
    n = 1000
    x, y = np.arange(n, dtype=np.float32), np.ones(n, np.float32)
    bx, by = upload(T, x), upload(T, y)
    a = ctypes.c_float(2.0)
    px, py, nn = ctypes.c_void_p(bx.ptr), ctypes.c_void_p(by.ptr), ctypes.c_int(n)
    args = (ctypes.c_void_p * 4)(ctypes.addressof(a), ctypes.addressof(px), ctypes.addressof(py), ctypes.addressof(nn))
    grid = (ctypes.c_uint32 * 3)(4, 1, 1)
    block = (ctypes.c_uint32 * 3)(256, 1, 1)
    capi.check(lib.cs_launch(T.stream, k, grid, block, 0, args, 4))
    assert np.array_equal(download(T, by, n), 2 * x + 1)
    bad = (ctypes.c_uint32 * 3)(2048, 1, 1)
    with pytest.raises(capi.InvalidWorkGroupSize):
        capi.check(lib.cs_launch(T.stream, k, grid, bad, 0, args, 4))
    capi.check(lib.cs_kernel_release(k))
    capi.check(lib.cs_program_release(prog))
    bx.release()
    by.release()


def test_build_failure_carries_log(gpu_tensors):
    prog, log = ctypes.c_void_p(), ctypes.c_void_p()
    st = lib.cs_program_build(gpu_tensors.device, b"extern \"C\" __global__ void k() { nope(); }", None,
                              ctypes.byref(prog), ctypes.byref(log))
    assert st == -11
    assert "nope" in capi.take_string(log.value)


def test_events_and_completion_callback(gpu_tensors):
    T = gpu_tensors
    e1, e2 = ctypes.c_void_p(), ctypes.c_void_p()
    capi.check(lib.cs_event_record(T.stream, ctypes.byref(e1)))
    b = alloc(T, 1 << 20)
    capi.check(lib.cs_random(T.stream, b.ptr, 1 << 20, 1))
    capi.check(lib.cs_event_record(T.stream, ctypes.byref(e2)))
    done = []
    CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
    cb = CB(lambda user: done.append(1))
    capi.check(lib.cs_event_on_complete(e2, ctypes.cast(cb, ctypes.c_void_p), None))
    capi.check(lib.cs_event_wait(e2))
    ms = ctypes.c_float()
    capi.check(lib.cs_event_elapsed_ms(e1, e2, ctypes.byref(ms)))
    assert ms.value >= 0
    c = ctypes.c_int()
    capi.check(lib.cs_event_query(e2, ctypes.byref(c)))
    assert c.value == 1
    import time
    for _ in range(200):
        if done:
            break
        time.sleep(0.01)
    assert done
    for e in (e1, e2):
        capi.check(lib.cs_event_release(e))
    b.release()


def test_graph_replays_fused_and_reduction(gpu_tensors):
    """A captured batch of fused launches and a ticketed reduction replays correctly
    several times, with eager work and synchronisation between replays."""
    T = gpu_tensors
    n = 1 << 18
    x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    tx = T.Tensor(x).doCache()
    e = T.Tensor.tanh(tx) * tx + tx
    f, leaves, _ = T.compile(e.closure, e.shape)
    out = alloc(T, n)
    s = alloc(T, 1)
    xb = tx.buffer
    leaf_ptrs = (ctypes.c_void_p * 1)(xb.ptr)
    capi.check(lib.cs_reduce_sum(T.stream, out.ptr, n, s.ptr, None))  # size the scratch outside capture
    T.synchronize()
    capi.check(lib.cs_graph_begin_capture(T.stream))
    capi.check(lib.cs_fused_launch(T.stream, f, leaf_ptrs, 1, out.ptr))
    capi.check(lib.cs_reduce_sum(T.stream, out.ptr, n, s.ptr, None))
    g = ctypes.c_void_p()
    capi.check(lib.cs_graph_end_capture(T.stream, ctypes.byref(g)))
    eager = e.flatArray()
    want_sum = eager.astype(np.float64).sum()
    for _ in range(4):
        capi.check(lib.cs_fill_uniform(T.stream, out.ptr, n, 5, -9.0, 9.0))  # clobber between replays
        capi.check(lib.cs_graph_launch(T.stream, g))
        T.synchronize()
        assert np.array_equal(download(T, out, n), eager)
        assert abs(float(download(T, s, 1)[0]) - want_sum) <= 4 * 2 ** -23 * np.abs(eager).astype(np.float64).sum()
    capi.check(lib.cs_graph_release(g))
    capi.check(lib.cs_fused_release(f))
    tx.release()
    out.release()
    s.release()


def test_graph_replays_after_the_scratch_grows(gpu_tensors):
    """A sum captured on a fresh stream uses that stream's small scratch; a larger sum run
    afterwards outside capture grows the scratch. The graph still addresses the old buffer,
    which must stay allocated: every replay gives the eager sum."""
    T = gpu_tensors
    s = ctypes.c_void_p()
    capi.check(lib.cs_stream_create(T.device, ctypes.byref(s)))
    small_n, big_n = 1 << 12, 1 << 26
    x = np.random.default_rng(2).standard_normal(big_n).astype(np.float32)
    xb = upload(T, x)
    out, out_big = alloc(T, 1), alloc(T, 1)
    T.synchronize()
    capi.check(lib.cs_reduce_sum(s, xb.ptr, small_n, out.ptr, None))  # sizes the scratch for small_n
    capi.check(lib.cs_stream_synchronize(s))
    capi.check(lib.cs_graph_begin_capture(s))
    capi.check(lib.cs_reduce_sum(s, xb.ptr, small_n, out.ptr, None))
    g = ctypes.c_void_p()
    capi.check(lib.cs_graph_end_capture(s, ctypes.byref(g)))
    capi.check(lib.cs_reduce_sum(s, xb.ptr, big_n, out_big.ptr, None))  # grows the scratch
    want = x[:small_n].astype(np.float64).sum()
    for _ in range(3):
        capi.check(lib.cs_fill_uniform(s, out.ptr, 1, 5, -9.0, 9.0))
        capi.check(lib.cs_graph_launch(s, g))
        capi.check(lib.cs_reduce_sum(s, xb.ptr, big_n, out_big.ptr, None))
        capi.check(lib.cs_stream_synchronize(s))
        got = float(download(T, out, 1)[0])
        assert abs(got - want) <= fp64.sum_bound(x[:small_n]), (got, want)
    assert abs(float(download(T, out_big, 1)[0]) - x.astype(np.float64).sum()) <= fp64.sum_bound(x)
    capi.check(lib.cs_graph_release(g))
    capi.check(lib.cs_stream_destroy(s))
    for h in (xb, out, out_big):
        h.release()


def test_release_during_capture(gpu_tensors):
    """A buffer released while a stream captures a graph: its release events skip the
    capturing stream (no event node enters the graph), later allocations that wait for
    pending releases succeed, and the graph replays."""
    T = gpu_tensors
    n = 1 << 16
    out = alloc(T, n)
    junk = [alloc(T, 1 << 20) for _ in range(4)]
    capi.check(lib.cs_graph_begin_capture(T.stream))
    capi.check(lib.cs_fill_uniform(T.stream, out.ptr, n, 9, -1.0, 1.0))
    for b in junk:
        b.release()  # released mid-capture
    g = ctypes.c_void_p()
    capi.check(lib.cs_graph_end_capture(T.stream, ctypes.byref(g)))
    big = [alloc(T, 64 << 20) for _ in range(4)]  # may wait on the pending releases
    for _ in range(3):
        capi.check(lib.cs_fill_uniform(T.stream, out.ptr, n, 1, 5.0, 6.0))
        capi.check(lib.cs_graph_launch(T.stream, g))
        T.synchronize()
        x = download(T, out, n)
        assert x.min() >= -1.0 and x.max() <= 1.0
    capi.check(lib.cs_graph_release(g))
    for b in big + [out]:
        b.release()


def test_release_then_stream_destroyed():
    """Buffers released while a second context's stream exists, that stream then destroyed:
    later allocations that settle the pending releases still succeed (no release event is
    waited for after the stream it was recorded on is gone)."""
    from compute import Tensors
    keep = Tensors(0)
    try:
        for round_ in range(6):
            with Tensors(0) as T:
                x = T.Tensor(np.full((256, 1024), float(round_), np.float32))
                y = (x + x).doCache()
                assert float(y.flatArray()[0]) == 2.0 * round_
                y.release()
            bufs = [alloc(keep, 8 << 20) for _ in range(3)]  # waits on the releases above
            for b in bufs:
                b.release()
    finally:
        keep.close()


def test_timed_launch_matches_and_times(gpu_tensors):
    """cs_fused_launch_timed computes what cs_fused_launch computes and returns events
    stamped by the dispatch: a positive duration below the host-side wall time."""
    import time
    T = gpu_tensors
    x = np.arange(1 << 20, dtype=np.float32)
    t = T.Tensor(x).doCache()
    e = t * t + t
    f, leaves, _ = T.compile(e.closure, e.shape)
    out1, out2 = alloc(T, x.size), alloc(T, x.size)
    ptrs = (ctypes.c_void_p * 1)(t.buffer.ptr)
    capi.check(lib.cs_fused_launch(T.stream, f, ptrs, 1, out1.ptr))
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    T.synchronize()
    w0 = time.perf_counter()
    capi.check(lib.cs_fused_launch_timed(T.stream, f, ptrs, 1, out2.ptr, ctypes.byref(e0), ctypes.byref(e1)))
    T.synchronize()
    wall_ms = (time.perf_counter() - w0) * 1e3
    ms = ctypes.c_float()
    capi.check(lib.cs_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
    assert 0.0 < ms.value < wall_ms
    assert np.array_equal(download(T, out1, x.size).view(np.uint32), download(T, out2, x.size).view(np.uint32))
    for ev in (e0, e1):
        capi.check(lib.cs_event_release(ev))
    lib.cs_fused_release(f)
    out1.release()
    out2.release()
    t.release()


def test_threads_share_one_device():
    """Eight host threads, each with its own Tensors (its own stream) on GPU 0, compile the
    same and different expressions at once (the structural cache's race path), allocate and
    release through the shared arena and read back: every result bit-exact."""
    import threading
    from compute import Tensors
    from oracle import inputs
    errors = []

    def work(i):
        try:
            with Tensors(0) as T:
                x = inputs.uniform(64 * 4096, 100 + i).reshape(64, 4096)
                for rep in range(6):
                    t = T.Tensor(x)
                    same = (t * t + t).flatArray()  # one structure in every thread
                    own = (t * T.Tensor.fill(float(i + rep), x.shape)).flatArray()  # a literal per thread
                    assert np.array_equal(same.view(np.uint32), ((x * x) + x).ravel().view(np.uint32))
                    assert np.array_equal(own.view(np.uint32), (x * np.float32(i + rep)).ravel().view(np.uint32))
        except Exception as ex:  # collected and asserted in the main thread
            errors.append((i, repr(ex)))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors
