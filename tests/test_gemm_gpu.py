"""matrixMultiply against the FP64 oracle and the reference's unrolled idioms."""
import numpy as np
import pytest

import spec_cases
from compute import _capi as capi
from gpu_util import alloc, download, upload
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu

VARIANTS = {"fp32": capi.GEMM_FP32, "bf16": capi.GEMM_BF16, "bf16x3": capi.GEMM_BF16X3, "fp8": capi.GEMM_FP8}
# Normalized error max |C - C64| / (|A| |B|) bounds: FP32 accumulation; BF16 rounds each
# operand to 8 bits (2 * 2^-9 per product); BF16x3 keeps about 16 bits (<= 1e-5 by design);
# e4m3 keeps 4 bits (2 * 2^-4 per product) with per-block power-of-two scales.
BOUND = {"fp32": 1e-6, "bf16": 2 * 2 ** -9 + 1e-6, "bf16x3": 1e-5, "fp8": 2 * 2 ** -4 + 1e-3}


@pytest.mark.parametrize("variant", ["fp32", "bf16", "bf16x3"])
def test_small_known_answer(gpu_tensors, variant):
    T = gpu_tensors
    c = T.Tensor.matrixMultiply(T.Tensor(spec_cases.M1), T.Tensor(spec_cases.M2), VARIANTS[variant])
    assert c.toString() == spec_cases.GOLDEN["matrix_multiplication"]["expect"]


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("mnk", [(1, 1, 1), (33, 65, 17), (128, 256, 512), (300, 200, 1000), (256, 512, 384),
                                 (513, 257, 129), (300, 200, 1024), (1, 65, 64)])
def test_gemm_shapes(gpu_tensors, variant, mnk):
    """Padded tiles, K tails, and row counts that end inside a tile."""
    T = gpu_tensors
    M, N, K = mnk
    a = inputs.uniform(M * K, 4001).reshape(M, K)
    b = inputs.uniform(K * N, 4002).reshape(K, N)
    ba, bb, bc = upload(T, a), upload(T, b), alloc(T, M * N)
    capi.check(capi.lib.cs_gemm(T.stream, VARIANTS[variant], M, N, K, ba.ptr, K, bb.ptr, N, bc.ptr, N))
    c = download(T, bc, M * N).reshape(M, N)
    err = fp64.gemm_normalized_error(c, a, b)
    print("%s %s normalized error %.3g" % (variant, mnk, err))
    assert err <= BOUND[variant]
    for h in (ba, bb, bc):
        h.release()


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_gemm_4096(gpu_tensors, variant):
    T = gpu_tensors
    n = 4096
    for scaled in (False, True):
        a, b = inputs.c4_inputs(n, scaled)
        ba, bb, bc = upload(T, a), upload(T, b), alloc(T, n * n)
        capi.check(capi.lib.cs_gemm(T.stream, VARIANTS[variant], n, n, n, ba.ptr, n, bb.ptr, n, bc.ptr, n))
        c = download(T, bc, n * n).reshape(n, n)
        rows = np.arange(3, n, 8)  # 512 rows, checked against FP64 in full
        c64 = a[rows].astype(np.float64) @ b.astype(np.float64)
        err = fp64.gemm_normalized_error(c[rows], a[rows], b, c64)
        print("%s 4096 scaled=%s normalized error %.3g" % (variant, scaled, err))
        assert err <= BOUND[variant]
        for h in (ba, bb, bc):
            h.release()


def test_fp8_representation(gpu_tensors):
    """OCP e4m3 (max 448, 3 mantissa bits): a matrix whose entries are exact e4m3 values
    times one power of two per row multiplies exactly."""
    T = gpu_tensors
    rng = np.random.default_rng(3)
    vals = np.array([0.0, 0.5, 1.0, 1.5, 2.0, 3.5, 6.0, 14.0, 448.0, -1.0, -3.0], np.float32)
    a = rng.choice(vals, size=(64, 128)).astype(np.float32)
    b = rng.choice(vals[:6], size=(128, 64)).astype(np.float32)  # all sums stay exact in FP32
    ba, bb, bc = upload(T, a), upload(T, b), alloc(T, 64 * 64)
    capi.check(capi.lib.cs_gemm(T.stream, capi.GEMM_FP8, 64, 64, 128, ba.ptr, 128, bb.ptr, 64, bc.ptr, 64))
    c = download(T, bc, 64 * 64).reshape(64, 64)
    assert np.array_equal(c, (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32))
    for h in (ba, bb, bc):
        h.release()


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_empty_inner_dimension(gpu_tensors, variant):
    """K = 0: every output is an empty sum, C = 0, and A and B are not read (null here)."""
    T = gpu_tensors
    M, N = 70, 300
    bc = upload(T, np.full(M * N, 7.0, np.float32))
    capi.check(capi.lib.cs_gemm(T.stream, VARIANTS[variant], M, N, 0, None, 0, None, N, bc.ptr, N))
    c = download(T, bc, M * N)
    assert np.array_equal(c, np.zeros(M * N, np.float32))
    bc.release()


def test_idioms_known_answer_both_routes():
    """TensorsSpec's two matrix-multiply idioms print [[74.0,...],[173.0,...]] through the
    reference's own kernels (small: below the thresholds) and through the GEMM route."""
    from compute import Tensors
    from test_gemm_idioms import mm1, mm2
    for route in ("large", "always"):
        with Tensors(0, gemm_idioms=route) as T:
            m1, m2 = T.Tensor(spec_cases.M1), T.Tensor(spec_cases.M2)
            before = T.idiom_gemms
            for f in (mm1, mm2):
                assert f(T, m1, m2).toString() == spec_cases.GOLDEN["matrix_multiplication"]["expect"], (route, f)
            assert T.idiom_gemms - before == (2 if route == "always" else 0)


def test_mm2_4096_is_one_gemm():
    """matrixMultiply2 written as the reference writes it, at 4096^3 (its [i,j,k] product
    would be 256 GiB): one cs_gemm call, 512 rows of C against FP64 within the FP32 bound."""
    from compute import Tensors
    from test_gemm_idioms import mm2
    n = 4096
    a, b = inputs.c4_inputs(n)
    with Tensors(0) as T:
        ta, tb = T.Tensor(a), T.Tensor(b)
        before = (T.launches, T.idiom_gemms)
        c = mm2(T, ta, tb).flatArray().reshape(n, n)
        assert (T.launches - before[0], T.idiom_gemms - before[1]) == (1, 1)
    rows = np.arange(5, n, 8)
    err = fp64.gemm_normalized_error(c[rows], a[rows], b)
    print("matrixMultiply2 at 4096^3 through the GEMM: normalized error %.3g" % err)
    assert err <= BOUND["fp32"]


def test_mm1_is_one_gemm():
    """matrixMultiply1 with 64 x 128 = 8192 unrolled products per output row: one cs_gemm
    call, within the FP32 bound of FP64; below the threshold (8 x 8) the reference's fused
    kernel, bit-exact with its FP32 left fold."""
    from compute import Tensors
    from test_gemm_idioms import mm1
    a = inputs.uniform(1024 * 64, 4011).reshape(1024, 64)
    b = inputs.uniform(64 * 128, 4012).reshape(64, 128)
    with Tensors(0) as T:
        before = T.idiom_gemms
        c = mm1(T, T.Tensor(a), T.Tensor(b)).flatArray().reshape(1024, 128)
        assert T.idiom_gemms - before == 1
        err = fp64.gemm_normalized_error(c, a, b)
        assert err <= BOUND["fp32"], err
        a8, b8 = a[:, :8].copy(), b[:8, :8].copy()
        c8 = mm1(T, T.Tensor(a8), T.Tensor(b8)).flatArray().reshape(1024, 8)
        assert T.idiom_gemms - before == 1
    want = np.zeros((1024, 8), np.float32)
    for j in range(8):
        want = (want + a8[:, j:j + 1] * b8[j:j + 1, :]) if j else a8[:, :1] * b8[:1, :]
    assert np.array_equal(c8.view(np.uint32), want.view(np.uint32))
