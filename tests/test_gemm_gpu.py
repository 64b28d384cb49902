"""matrixMultiply against the FP64 oracle and the reference's unrolled idioms."""
import numpy as np
import pytest

import spec_cases
from compute import _capi as capi
from gpu_util import alloc, download, upload
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu

VARIANTS = {"fp32": capi.GEMM_FP32}
BOUND = {"fp32": 1e-6}


def test_small_known_answer(gpu_tensors):
    T = gpu_tensors
    c = T.Tensor.matrixMultiply(T.Tensor(spec_cases.M1), T.Tensor(spec_cases.M2))
    assert c.toString() == spec_cases.GOLDEN["matrix_multiplication"]["expect"]


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("mnk", [(1, 1, 1), (33, 65, 17), (128, 256, 512), (300, 200, 1000)])
def test_gemm_shapes(gpu_tensors, variant, mnk):
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
        rows = np.arange(0, n, 64)
        c64 = a[rows].astype(np.float64) @ b.astype(np.float64)
        err = fp64.gemm_normalized_error(c[rows], a[rows], b, c64)
        print("%s 4096 scaled=%s normalized error %.3g" % (variant, scaled, err))
        assert err <= BOUND[variant]
        for h in (ba, bb, bc):
            h.release()
