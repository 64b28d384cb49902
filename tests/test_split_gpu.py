"""Fusion splits of a join: cutting a join into g kernels (plus the kernel that gathers
their outputs) gives the one-kernel join's bits, and the split cost model's cut is within 10%
of the exhaustive autotuner's best on joins it never saw."""
import ctypes

import numpy as np
import pytest

from compute import _capi as capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("groups", [2, 4, 8])
def test_split_join_is_bit_identical(groups):
    """The convolution idiom (8 filters, 3x3 kernel, 4 channels) materialised as one kernel
    and cut into 2, 4 and 8 kernels: identical bits."""
    from compute import Tensors, workloads
    from oracle import inputs
    x = inputs.uniform(2 * 16 * 16 * 4, 811).reshape(2, 16, 16, 4)
    w = inputs.uniform(3 * 3 * 4 * 8, 812).reshape(3, 3, 4, 8)
    b = inputs.uniform(8, 813)
    outs = {}
    for g in (1, groups):
        with Tensors(0, vector_width=1, join_split=g) as T:
            y = workloads.convolute(T, T.Tensor(x), T.Tensor(w), T.Tensor(b))
            before = T.launches
            outs[g] = y.flatArray()
            assert T.launches - before == (1 if g == 1 else g + 1)
    assert np.array_equal(outs[1].view(np.uint32), outs[groups].view(np.uint32))


def test_split_model_on_new_joins():
    """The shipped split model (trained on the exhaustive sweep of the first 192 of 256 joins of
    seed 6001) picks a cut for 12 joins of seed 6101 that it never saw; every cut of every join
    is timed live:
    the total time of its picks is within 10% of the exhaustive best."""
    from compute import Tensors, autotune, costmodel as cm, workloads
    with Tensors(0) as T:
        model = cm.load_split_model(T)
        assert model is not None
        joins = workloads.conv_joins(T, 12, 6101)
        bufs, ptrs = [], {}
        for j in joins:
            for t, d in zip(j.leaves, j.leaf_data):
                a = np.ascontiguousarray(d, np.float32).ravel()
                buf = T._alloc(a.nbytes)
                capi.check(capi.lib.cs_buffer_write(T.stream, buf.handle, 0, a.ctypes.data_as(ctypes.c_void_p),
                                                    a.nbytes, 1))
                bufs.append(buf)
                ptrs[id(t)] = buf.ptr
        best = chosen = never = 0.0
        for j in joins:
            times = {g: autotune.time_join(T, j.tensor, ptrs, g, reps=20) for g in autotune.split_candidates(j.parts)}
            pick = model.choose_groups(j.tensor)
            best += min(times.values())
            chosen += times[pick]
            never += times[1]
            print("join %s parts %d: times %s, model %d" % (j.desc, j.parts, {g: round(t, 1) for g, t in times.items()},
                                                            pick), flush=True)
        model.release()
        for buf in bufs:
            buf.release()
        print("exhaustive %.1f us, model %.1f us (regret %.3f), one kernel %.1f us" %
              (best, chosen, chosen / best - 1.0, never))
        assert chosen <= 1.10 * best
