"""The persistent batch of uneven expressions: bit-identical to per-expression launches,
within the propagated error bound of the oracle, and correct on every graph replay."""
import numpy as np
import pytest

from compute import batch as bm
from compute import workloads
from oracle import fp64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_set(gpu_tensors):
    T = gpu_tensors
    exprs = workloads.uneven_dags(T, 48, seed=77, max_elements=200_000)
    eset = bm.ExpressionSet(T, exprs)
    yield eset
    eset.release()


def test_batch_matches_per_expression_launches(small_set):
    eset = small_set
    T = eset.T
    rng = np.random.default_rng(0)
    # non-temporal stores on some bodies: a run of 4 (or 1) floats writes through at device
    # scope, through the batch's global-address-space pointers
    cfgs = [{"vector_width": int(rng.choice([1, 2, 4])), "items_per_thread": int(rng.choice([1, 2, 4])),
             "block_size": 256, "nontemporal_store": int(rng.integers(0, 2))} for _ in eset.exprs]
    fused = bm.compile_many(T, eset.exprs, cfgs)
    for i, f in enumerate(fused):
        eset.launch(i, f, int(rng.choice([64, 256, 1024])))
    each = eset.outputs()
    eset.clear_outputs()
    b = bm.Batch(eset, fused, block=256)
    g = bm.Graph(T, b.launch)
    for rep in range(3):
        eset.clear_outputs()
        g.launch()
        T.synchronize()
        got = eset.outputs()
        for i in range(len(each)):
            assert np.array_equal(got[i].view(np.uint32), each[i].view(np.uint32)), (rep, i)
    g.release()
    b.release()
    from compute import _capi as capi
    for f in fused:
        capi.lib.cs_fused_release(f)
    # and the results are right: within the propagated bound of the FP32 program
    for i in range(0, len(each), 6):
        e = eset.exprs[i]
        value, bound = fp64.evaluate_with_bound(e.tensor.closure, e.shape, e.leaf_data, ulps=2.0)
        assert fp64.within_bound(each[i], value, bound).all(), i


def test_batch_of_one_and_tiny(gpu_tensors):
    T = gpu_tensors
    exprs = workloads.uneven_dags(T, 3, seed=5, min_elements=1, max_elements=70)
    eset = bm.ExpressionSet(T, exprs)
    try:
        fused = bm.compile_many(T, exprs, [bm.DEFAULT] * len(exprs))
        b = bm.Batch(eset, fused, block=64)
        b.launch()
        T.synchronize()
        got = eset.outputs()
        for i, e in enumerate(exprs):
            value, bound = fp64.evaluate_with_bound(e.tensor.closure, e.shape, e.leaf_data, ulps=2.0)
            assert fp64.within_bound(got[i], value, bound).all()
        b.release()
        from compute import _capi as capi
        for f in fused:
            capi.lib.cs_fused_release(f)
    finally:
        eset.release()


def test_c5_full_size(gpu_tensors):
    """The batched workload at its stated size: 512 expressions of seed 5001 (3-12 ops, 1e3 to
    4e6 elements, permute and broadcast views), bodies V=2, U=2 in 512-thread workgroups. The
    graph-captured persistent batch gives the same bits as per-expression launches of the
    same kernels for all 512 outputs, and every 16th output is within the propagated bound of
    the FP32 program with correctly rounded operations."""
    from compute import _capi as capi
    T = gpu_tensors
    exprs = workloads.uneven_dags(T, 512, 5001)
    print("c5: %d expressions, %.2f GB" % (len(exprs), sum(e.bytes for e in exprs) / 1e9), flush=True)
    eset = bm.ExpressionSet(T, exprs)
    try:
        cfg = {"vector_width": 2, "items_per_thread": 2, "block_size": 512, "nontemporal_store": 1}
        fused = bm.compile_many(T, exprs, [cfg] * len(exprs))
        for i, f in enumerate(fused):
            eset.launch(i, f, 512)
        each = eset.outputs()
        eset.clear_outputs()
        b = bm.Batch(eset, fused, block=512)
        g = bm.Graph(T, b.launch)
        eset.clear_outputs()
        g.launch()
        T.synchronize()
        got = eset.outputs()
        same = [np.array_equal(got[i].view(np.uint32), each[i].view(np.uint32)) for i in range(len(exprs))]
        assert all(same), [i for i, s in enumerate(same) if not s][:10]
        g.release()
        b.release()
        for f in fused:
            capi.lib.cs_fused_release(f)
        for i in range(0, len(exprs), 16):
            e = exprs[i]
            value, bound = fp64.evaluate_with_bound(e.tensor.closure, e.shape, e.leaf_data, ulps=2.0)
            assert fp64.within_bound(each[i], value, bound).all(), i
    finally:
        eset.release()


def test_row_groups_in_a_batch(gpu_tensors):
    """Bodies at one element per vector whose gathers read a pixel's channels (split along the
    last dimension, some translated to the edges) share one row per pixel in the batch's
    device functions as in stand-alone kernels: the batch is bit-identical to per-expression
    launches and within the oracle's bound."""
    from compute import _capi as capi
    from oracle import inputs
    T = gpu_tensors
    exprs = []
    for k, (h, w, d) in enumerate([(5, 9, 8), (17, 3, 12), (4, 33, 6), (64, 70, 4)]):
        x = inputs.uniform(h * w * d, 300 + k).reshape(h, w, d)
        tx = T.Tensor(x)
        c = list(tx.split(2))
        t = c[0] * c[1] + c[d - 1] - c[2].translate([1, -1]) * c[d // 2]
        exprs.append(workloads.Expr(k, t, [tx], [x], [], ["direct"]))
    eset = bm.ExpressionSet(T, exprs)
    try:
        cfg = {"vector_width": 1, "items_per_thread": 2, "block_size": 256, "nontemporal_store": 0}
        fused = bm.compile_many(T, exprs, [cfg] * len(exprs))
        for i, f in enumerate(fused):
            eset.launch(i, f, 256)
        each = eset.outputs()
        eset.clear_outputs()
        b = bm.Batch(eset, fused, block=256)
        b.launch()
        T.synchronize()
        got = eset.outputs()
        for i, e in enumerate(exprs):
            assert np.array_equal(got[i].view(np.uint32), each[i].view(np.uint32)), i
            value, bound = fp64.evaluate_with_bound(e.tensor.closure, e.shape, e.leaf_data, ulps=2.0)
            assert fp64.within_bound(got[i], value, bound).all(), i
        b.release()
        for f in fused:
            capi.lib.cs_fused_release(f)
    finally:
        eset.release()
