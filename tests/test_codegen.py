"""HIP code generation without a device: the emitted text folds views into index math,
compiles for gfx950 with hiprtc, and rejects malformed DAGs (CPU)."""
import ctypes

import numpy as np
import pytest

from compute import _capi as capi
from compute import trees


def compiles(src, opts=b"-ffp-contract=off"):
    log = ctypes.c_void_p()
    st = capi.lib.cs_program_check(src.encode(), opts, ctypes.byref(log))
    text = capi.take_string(log.value)
    assert st == 0, text
    return True


def c2_expr(T):
    X = T.Tensor(np.zeros((64, 256, 256), np.float32))
    bias = T.Tensor(np.zeros((256,), np.float32))
    p = X.permute([1, 0, 2])
    y = p + bias.broadcast((256, 64, 256)) + p.translate([0, 0, 1])
    return T.Tensor.exp(-(y * y)) * T.Tensor.tanh(y)


def test_c1_kernel_text(offline_tensors):
    T = offline_tensors
    a, b, c = (T.Tensor(np.zeros((1024, 1024), np.float32)) for _ in range(3))
    r = a * b + c
    src = T.kernel_source(r.closure, r.shape)
    assert src.count("const csmi_f4 a0_") == 3  # three wide loads
    assert "t4[u][v] = t1[u][v] * t3[u][v]" in src and "t7[u][v] = t4[u][v] + t6[u][v]" in src
    body = src[src.index("void csmi_fused_"):]
    assert "fma" not in body  # no contraction in the precise mode
    compiles(src)


def test_c2_views_fold_into_one_kernel(offline_tensors):
    T = offline_tensors
    out = c2_expr(T)
    nodes, outputs, leaves = trees.serialize(out.closure)
    assert len(leaves) == 2  # X once (both views share its parameter) and bias
    src = T.kernel_source(out.closure, out.shape)
    assert "p0 + (g2_i0 * 65536 + g2_i1 * 256 + g2_i2)" in src  # permute(1,0,2)
    assert "(int)(c2[u] + -1)" in src  # translate(0,0,1)
    assert "p1[g5_i0]" in src  # bias aligned with dimension 0, one scalar load
    # X and its translate share the aligned chunk at the vector's own position: two 16-byte
    # loads of X per vector of four
    assert src.count("l = *(const csmi_f4*)") == 2 and "t8[u][1] = g2_w0[u][0];" in src
    assert "csmi_expf_np(" in src  # exp(-(y*y)) cannot overflow
    assert "csmi_tanhf_p(" in src
    compiles(src)
    from compute import Tensors
    acc = Tensors(offline=True, math_mode=capi.MATH_ACCURATE)
    src = acc.kernel_source(c2_expr(acc).closure, (256, 64, 256))
    assert "csmi_expf(" in src and "csmi_tanhf(" in src
    compiles(src)


def test_launch_order_tiles_across_a_transpose(offline_tensors):
    """A leaf read transposed (its rows run along output dimension 0) makes the kernel take
    tiles across that dimension: 32 positions by 16 vectors for [96, 64], the whole extents
    for [7, 9]. C2's permute keeps X's rows along the innermost dimension and stays linear,
    as do a transpose whose rows are too short (2 elements) and the linear launch order."""
    from compute import Tensors
    T = offline_tensors
    x = T.Tensor(np.zeros((64, 96), np.float32))
    y = T.Tensor(np.zeros((96, 64), np.float32))
    r = x.permute([1, 0]) * T.Tensor.fill(2.0, (96, 64)) + y
    src = T.kernel_source(r.closure, r.shape)
    assert "% 512u" in src and "dt * 32u" in src  # R = 32 positions of dimension 0, C = 16 vectors
    compiles(src)
    assert "dt * " not in T.kernel_source(c2_expr(T).closure, (256, 64, 256))
    small = T.Tensor(np.zeros((9, 7), np.float32)).permute([1, 0]) + T.Tensor.fill(1.0, (7, 9))
    src = T.kernel_source(small.closure, small.shape)
    assert "dt * 7u" in src  # R = 7: the whole extent
    compiles(src)
    short = T.Tensor(np.zeros((64, 2), np.float32)).permute([1, 0]) + T.Tensor.fill(1.0, (2, 64))
    assert "dt * " not in T.kernel_source(short.closure, short.shape)
    lin = Tensors(offline=True, launch_order=capi.ORDER_LINEAR)
    x2 = lin.Tensor(np.zeros((64, 96), np.float32))
    r2 = x2.permute([1, 0]) * lin.Tensor.fill(2.0, (96, 64))
    assert "dt * " not in lin.kernel_source(r2.closure, r2.shape)
    x3 = T.Tensor(np.zeros((4, 64, 96), np.float32))
    r3 = x3.permute([2, 0, 1]) + T.Tensor.fill(1.0, (96, 4, 64))  # X's rows run along dimension 0
    src3 = T.kernel_source(r3.closure, r3.shape)
    assert "dt * 32u" in src3
    compiles(src3)


@pytest.mark.parametrize("shape", [(), (7,), (3, 5), (2, 3, 4, 5), (2, 1, 3, 1, 2, 2)])
def test_shapes_compile(offline_tensors, shape):
    T = offline_tensors
    x = T.Tensor(np.zeros(shape, np.float32))
    r = T.Tensor.max(x + T.Tensor.fill(1.0, shape), x * x) - T.Tensor.sqrt(T.Tensor.abs(x))
    compiles(T.kernel_source(r.closure, r.shape))


def test_fractional_scale_and_tuple_compile(offline_tensors):
    T = offline_tensors
    x = T.Tensor(np.zeros((4, 6), np.float32))
    s = x.scale((6, 9))
    src = T.kernel_source(s.closure, s.shape)
    assert "0.667" in src  # coefficient printed with three decimals, as the reference does
    compiles(src)
    j = T.Tensor.join([x, x + x, x * x])
    src = T.kernel_source(trees.Concatenate([p.closure for p in j._parts]), (4, 6))
    compiles(src)


def test_fast_math_mode_compiles(offline_tensors):
    from compute import Tensors
    T = Tensors(offline=True, math_mode=capi.MATH_FAST)
    out = c2_expr(T)
    src = T.kernel_source(out.closure, out.shape)
    assert "__expf(" in src
    compiles(src, b"-ffast-math -ffp-contract=fast")


def test_infinite_literal_and_mod_compile(offline_tensors):
    T = offline_tensors
    x = T.Tensor(np.zeros((8,), np.float32))
    r = T.Tensor.min(x % T.Tensor.fill(3.0, (8,)), T.Tensor.fill(float("inf"), (8,)))
    src = T.kernel_source(r.closure, r.shape)
    assert "fmodf(" in src and "__builtin_huge_valf()" in src
    compiles(src)


def _gen(nodes, outputs, shape):
    outs = (ctypes.c_int32 * len(outputs))(*outputs)
    shp = (ctypes.c_int32 * max(len(shape), 1))(*shape)
    src = ctypes.c_void_p()
    st = capi.lib.cs_fused_generate_source(nodes, len(nodes), outs, len(outputs), len(shape), shp, None,
                                           ctypes.byref(src), None)
    capi.take_string(src.value)
    return st


def test_malformed_dags_are_rejected():
    nodes = (capi.CsNode * 2)()
    nodes[0].op = capi.OP_PARAMETER
    nodes[0].param = 1  # must be 0: leaves are numbered in first-visit order
    nodes[0].rank = 1
    nodes[0].shape[0] = 4
    nodes[1].op = capi.OP_EXTRACT
    nodes[1].a = 0
    assert _gen(nodes, [1], [4]) == -30
    nodes[0].param = 0
    assert _gen(nodes, [1], [4]) == 0
    assert _gen(nodes, [0], [4]) == -30  # an array is not a float output
    nodes[1].a = 5
    assert _gen(nodes, [1], [4]) == -30  # operand out of range
    nodes[1].a = 1
    assert _gen(nodes, [1], [4]) == -30  # cycle
    nodes[1].a = 0
    assert _gen(nodes, [1], [1 << 16, 1 << 16]) == -61  # index math is 32-bit
    nodes[0].rank = 3
    nodes[0].shape[1] = nodes[0].shape[2] = 4
    assert _gen(nodes, [1], [4]) == -30  # a rank-3 leaf cannot be read in a rank-1 kernel


def test_structural_key_ignores_leaf_identity(offline_tensors):
    T = offline_tensors

    def key(expr):
        nodes, outputs, _ = trees.serialize(expr.closure)
        outs = (ctypes.c_int32 * len(outputs))(*outputs)
        shp = (ctypes.c_int32 * len(expr.shape))(*expr.shape)
        src = ctypes.c_void_p()
        k = ctypes.c_uint64()
        assert capi.lib.cs_fused_generate_source(nodes, len(nodes), outs, len(outputs), len(expr.shape), shp,
                                                 None, ctypes.byref(src), ctypes.byref(k)) == 0
        capi.take_string(src.value)
        return k.value

    a, b = T.Tensor(np.zeros((4, 4), np.float32)), T.Tensor(np.ones((4, 4), np.float32))
    assert key(a * b + a) == key(b * a + b)
    assert key(a * b + a) != key(a * b + b)
    assert key(a + b) != key(a.transpose() + b)


def test_batch_source_offline(offline_tensors):
    """The persistent batch's source (cs_batch_generate_source): one non-inlined device
    function per expression over global-address-space pointers and a run of tiles, a
    constant workgroup size, and one switch case per expression; the text compiles for
    gfx950."""
    from compute import batch as bm
    from compute import workloads
    T = offline_tensors
    exprs = workloads.uneven_dags(T, n=6, seed=11, min_elements=100, max_elements=5000)
    cfgs = [{"vector_width": 2, "items_per_thread": 2, "block_size": 512}] * len(exprs)
    descs, keep = bm._descs(exprs, cfgs, capi.MATH_PRECISE)
    src, items = ctypes.c_void_p(), ctypes.c_int64()
    capi.check(capi.lib.cs_batch_generate_source(descs, len(exprs), 512, ctypes.byref(src), ctypes.byref(items)))
    text = capi.take_string(src.value)
    assert text.count("__device__ __noinline__ void expr_") == len(exprs)
    assert "#define CSMI_G __attribute__((address_space(1)))" in text
    assert "(tile) * (512u * 2u)" in text and "blockDim" not in text.split("batch_kernel")[0]
    for i, e in enumerate(exprs):
        assert "case %d: expr_%d(it.y, it.y + it.z" % (i, i) in text
    # one call per work item: the run of tiles is looped inside the body's function
    assert text.count("for (unsigned tile = tile_begin; tile < tile_end; ++tile)") == len(exprs)
    # every tile of every expression is in exactly one work item
    tiles = sum(-(-(e.out_elements // bm.effective_vector_width(e.shape, 2)) // (512 * 2)) for e in exprs)
    assert 1 <= items.value <= tiles
    compiles(text)
    with pytest.raises(capi.InvalidWorkGroupSize):
        capi.check(capi.lib.cs_batch_generate_source(descs, len(exprs), 96, ctypes.byref(src), None))


@pytest.mark.parametrize("v", [1, 2, 4, 8])
def test_nontemporal_stores_write_through(v):
    """With non-temporal stores, a run of 4 floats per vector (one 16-byte store covering the
    wave's lanes contiguously) writes through at device scope (`sc1 nt`); longer and odd runs
    stay plain non-temporal stores (written through they would be partial-line writes)."""
    from compute import Tensors
    T = Tensors(offline=True, vector_width=v)
    T.options.nontemporal_store = 1
    x = T.Tensor(np.zeros((3, 7, 8), np.float32))
    j = T.Tensor.join([x + x, x * x])  # two outputs: runs of 2V floats
    src = T.kernel_source(trees.Concatenate([p.closure for p in j._parts]), (3, 7, 8))
    run = 2 * v
    assert src.count("off sc1 nt") == (1 if run == 4 else 0)
    assert src.count("__builtin_nontemporal_store(") == (0 if run == 4 else (run // 4 if run % 4 == 0 else run))
    compiles(src)
    y = T.Tensor(np.zeros((3, 7, 8), np.float32))
    src = T.kernel_source((y * y).closure, (3, 7, 8))  # one output: a run of V floats
    assert ("off sc1 nt" in src) == (v in (1, 4))
    compiles(src)


def test_leaves_of_2_31_elements_are_rejected(offline_tensors):
    """Gather offsets are 32-bit (the reference's int indices): slice 2 of split(0) on a
    [3, 1e9] leaf would address element 2e9 + i, so compiling it fails with
    CS_INVALID_BUFFER_SIZE instead of emitting an offset that wraps. A leaf just under 2^31
    elements still compiles."""
    from compute import tensors
    T = offline_tensors
    big = trees.ArrayParameter(object(), 0.0, (3, 1000000000))
    view = trees.Extract(trees.Transform(big, tensors._split_matrix(2, 0, 2), 1))
    with pytest.raises(capi.InvalidBufferSize):
        T.kernel_source(view, (1000000000,))
    ok = trees.ArrayParameter(object(), 0.0, ((1 << 31) - 1,))
    src = T.kernel_source(trees.Binary("+", trees.Extract(ok), trees.Extract(ok)), ((1 << 31) - 1,))
    assert "p0" in src


@pytest.mark.parametrize("D,F,V", [(3, 4, 1), (8, 4, 1), (12, 2, 1), (1, 2, 1), (8, 4, 4)])
def test_constant_table_and_row_groups(D, F, V):
    """The convolution idiom's body (workloads.convolute): 32 or more constant gathers (weights
    and biases picked out by split) are copied into an LDS table through a (slot, offset) table
    per leaf, in the order the gathers read them, and read from it; at one element per vector
    the channels of one input pixel share one bounds check and load in 16-byte chunks when the
    channel count is a multiple of 4. Every variant compiles for gfx950."""
    import re

    from compute import Tensors, workloads
    T = Tensors(offline=True, vector_width=V, items_per_thread=1, block_size=256)
    B, H, W = 2, 9, 12
    x = T.Tensor(np.zeros((B, H, W, D), np.float32))
    w = T.Tensor(np.zeros((3, 3, D, F), np.float32))
    b = T.Tensor(np.zeros((F,), np.float32))
    out = workloads.convolute(T, x, w, b)
    src = T.kernel_source(trees.Concatenate([p.closure for p in out._parts]), (B, H, W))
    staged = 9 * D * F + F
    assert ("__shared__ float csmi_k[%d];" % staged in src) == (staged >= 32)
    if staged >= 32:
        # the slots run 0 .. staged-1 in the order the gathers read them
        slots = sorted(int(k) for k in re.findall(r"\? csmi_k\[(\d+)\]", src))
        assert slots == list(range(staged))
        first = [int(k) for k in re.findall(r"\? csmi_k\[(\d+)\]", src)]
        assert first == sorted(first)
    groups = re.findall(r"\br(\d+)_ok = live\[u\]", src)
    # one per kernel offset for the input's channels; without the table, the weights of the
    # F filters at each (offset, channel) are one row of the weight leaf as well
    inputs_rows = 9 if D > 1 else 0
    weight_rows = 9 * D if staged < 32 and F > 1 else 0
    assert len(groups) == (inputs_rows + weight_rows if V == 1 else 0)
    chunks = re.findall(r"\br\d+_c(\d+) = \*\(const csmi_f4\*\)", src)
    assert len(chunks) == (9 * D // 4 if V == 1 and D % 4 == 0 else 0)
    compiles(src)


def test_register_budget_of_a_small_grid():
    """A body of more than 256 nodes whose grid gives few waves per SIMD of 256 CUs (4 SIMDs
    each) is compiled with the budget of that many waves, at least 2: 2048 waves and fewer get
    amdgpu_waves_per_eu(1, 2), 3072 waves (1, 3); a grid of 8 or more waves per SIMD gets none."""
    from compute import Tensors, workloads
    T = Tensors(offline=True, vector_width=1, items_per_thread=1, block_size=256)
    for batch, hint in ((32, 2), (128, 2), (192, 3), (1024, None)):
        x = T.Tensor(np.zeros((batch, 32, 32, 8), np.float32))
        w = T.Tensor(np.zeros((3, 3, 8, 8), np.float32))
        b = T.Tensor(np.zeros((8,), np.float32))
        out = workloads.convolute(T, x, w, b)
        src = T.kernel_source(trees.Concatenate([p.closure for p in out._parts]), (batch, 32, 32))
        if hint is None:
            assert "amdgpu_waves_per_eu" not in src
        else:
            assert "__attribute__((amdgpu_waves_per_eu(1, %d)))" % hint in src


def test_row_groups_in_batch_bodies(offline_tensors):
    """Rows read together also in a persistent batch's device functions (global-address-space
    row pointers); the batch source compiles for gfx950."""
    import re

    from compute import batch as bm
    T = offline_tensors

    class Expr(object):
        pass

    exprs = []
    for k in range(3):
        c = list(T.Tensor(np.zeros((6, 7 + k, 8), np.float32)).split(2))
        e = Expr()
        e.tensor = c[0] * c[1] + c[2] - c[5] * c[7] + c[3].translate([1, 0])
        e.shape = tuple(e.tensor.shape)
        exprs.append(e)
    descs, keep = bm._descs(exprs, [{"vector_width": 1, "items_per_thread": 2, "block_size": 256}] * 3,
                            capi.MATH_PRECISE)
    src = ctypes.c_void_p()
    capi.check(capi.lib.cs_batch_generate_source(descs, 3, 256, ctypes.byref(src), None))
    text = capi.take_string(src.value)
    assert len(re.findall(r"\br\d+_ok = live\[u\]", text)) == 3  # c0 c1 c2 c5 c7 of each body
    assert "const CSMI_G float* r" in text and "csmi_k" not in text
    compiles(text)


def test_large_bodies_default_to_one_element_per_thread(offline_tensors):
    """Without an explicit launch shape a body of more than 256 canonical nodes (the
    convolution idiom) runs one element per vector and per thread; a small body keeps 4
    elements per vector and 2 vectors per thread from 2^22 elements."""
    from compute import workloads
    T = offline_tensors
    x = T.Tensor(np.zeros((128, 32, 32, 8), np.float32))
    w = T.Tensor(np.zeros((3, 3, 8, 8), np.float32))
    b = T.Tensor(np.zeros((8,), np.float32))
    out = workloads.convolute(T, x, w, b)
    clo = trees.Concatenate([p.closure for p in out._parts])
    src = T.kernel_source(clo, (128, 32, 32))
    assert ", V=1 U=1\n" in src.split("\n", 1)[0] + "\n"
    y = T.Tensor(np.zeros((1024, 4096), np.float32))
    src = T.kernel_source((y * y).closure, (1024, 4096))
    assert ", V=4 U=2\n" in src.split("\n", 1)[0] + "\n"
