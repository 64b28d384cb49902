"""RNG kernels: `random` bit-exact with the reference program; `randomNormal` within 2 ulp
of it (libm differs); the synthetic-input generators equal the numpy definitions."""
import numpy as np
import pytest

import oracle
from compute import _capi as capi
from gpu_util import alloc, download
from oracle import fp64, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed", [(9, 12345), (1, 0), (1000003, 7)])
def test_random_bit_exact(gpu_tensors, n, seed):
    got = gpu_tensors.Tensor.random((n,), seed=seed).flatArray()
    assert np.array_equal(got.view(np.uint32), oracle.random(n, seed).view(np.uint32))


@pytest.mark.parametrize("n,seed", [(27, 54321), (1, 54321), (999999, 3)])
def test_random_normal_ulp(gpu_tensors, n, seed):
    got = gpu_tensors.Tensor.randomNormal((n,), seed=seed).flatArray()
    want = oracle.random_normal(n, seed)
    d = fp64.ulp_distance(got, want)
    finite = np.isfinite(want)
    print("randomNormal max ulp vs glibc float program: %d" % d[finite].max())
    assert d[finite].max() <= 2 or np.abs(got[finite] - want[finite]).max() <= 4e-7
    # where the host program gives +-inf or NaN (a uniform draw of exactly 0), so does the device
    assert np.array_equal(got[~finite], want[~finite], equal_nan=True)


def test_fill_generators_match_numpy(gpu_tensors):
    T = gpu_tensors
    n = 1 << 20
    b = alloc(T, n)
    capi.check(capi.lib.cs_fill_uniform(T.stream, b.ptr, n, 1001, -1.0, 1.0))
    assert np.array_equal(download(T, b, n), inputs.uniform(n, 1001))
    capi.check(capi.lib.cs_fill_normal(T.stream, b.ptr, n, 2001))
    d = fp64.ulp_distance(download(T, b, n), inputs.normal(n, 2001))
    assert d.max() <= 1
    b.release()
