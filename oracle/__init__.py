"""The oracle: CPU restatements of the reference's algorithms for the fused-expression,
reduction, RNG and matrix-multiply path.

TEST INFRASTRUCTURE. Only tests/, `__graft_entry__.smoke()` and the `cpu_baseline` leg
of `bench.py` may import, call, link or execute anything here, and only as the checker
(or, for the CPU baseline, as the timed reference CPU path). The product never routes
through it.

  refcl.py              -- the reference's generated OpenCL C kernels (OpenCLKernelBuilder
                           templates) emitted and run as host C (stands in for POCL)
  reference_kernels.c   -- the reference's fixed programs: sequential and parallel `sum`,
                           `random`, `random_normal`; CPU baselines (row folds,
                           sequential-k GEMM) and FP64 row statistics
  fp64.py               -- numpy interpreter of a closure DAG in FP64 / per-op FP32,
                           ulp and reduction-bound metrics, FP64 GEMM error
  inputs.py             -- Wang-hash seeded inputs

Pinning (every function is checked against something it does not compute itself):
  refcl kernels         -- TensorsSpec/cpuSpec/README known answers (tests/golden), and
                           cross-checked against fp64.evaluate on random DAGs
  refcl_sequential_sum  -- fill(15625,[8,8]).sum == 1000000.0 (TensorsSpec.scala:251-257),
                           FP64 sums within the FP32 error bound
  refcl_parallel_sum    -- power-of-two sizes against FP64; the dropped-element quirk
                           for non power-of-two local sizes against its closed form
  refcl_random          -- the 9 known values of TensorsSpec.scala:402-409, bit exact
  refcl_random_normal   -- the 27 known values of TensorsSpec.scala:411-434 within 2 ulp
  fp64.evaluate         -- closed forms (transpose, broadcast, translate with padding)
                           and the known answers
  inputs                -- wang_hash against refcl_random's known values
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def lib():
    """The host-C library, built on first use when absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build()
        l = ctypes.CDLL(LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        l.refcl_sequential_sum.restype = ctypes.c_float
        l.refcl_sequential_sum.argtypes = [fp, ctypes.c_uint32]
        l.refcl_parallel_sum.restype = ctypes.c_float
        l.refcl_parallel_sum.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        l.refcl_random.restype = None
        l.refcl_random.argtypes = [fp, ctypes.c_int64, ctypes.c_uint32]
        l.refcl_random_normal.restype = None
        l.refcl_random_normal.argtypes = [fp, ctypes.c_int64, ctypes.c_uint32]
        l.base_row_fold.restype = None
        l.base_row_fold.argtypes = [fp, ctypes.c_int64, ctypes.c_int64, fp, fp]
        l.base_gemm_seqk.restype = None
        l.base_gemm_seqk.argtypes = [fp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        l.fp64_row_stats.restype = None
        l.fp64_row_stats.argtypes = [fp, ctypes.c_int64, ctypes.c_int64, dp, fp, dp, dp]
        _lib = l
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def sequential_sum(x):
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    return np.float32(lib().refcl_sequential_sum(_fp(x), x.size))


def parallel_sum(x, work_group_size=256, compute_units=256):
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    scratch = np.zeros(max(compute_units, 1), dtype=np.float32)
    return np.float32(lib().refcl_parallel_sum(_fp(x), x.size, work_group_size, compute_units, _fp(scratch)))


def random(n, seed):
    out = np.zeros(n, dtype=np.float32)
    lib().refcl_random(_fp(out), n, seed & 0xFFFFFFFF)
    return out


def random_normal(n, seed):
    out = np.zeros(n + (n & 1), dtype=np.float32)
    lib().refcl_random_normal(_fp(out), n, seed & 0xFFFFFFFF)
    return out[:n]


def row_fold(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    rows, cols = a.shape
    s = np.zeros(rows, np.float32)
    m = np.zeros(rows, np.float32)
    lib().base_row_fold(_fp(a), rows, cols, _fp(s), _fp(m))
    return s, m


def gemm_seqk(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    M, K = a.shape
    N = b.shape[1]
    c = np.zeros((M, N), np.float32)
    lib().base_gemm_seqk(_fp(a), _fp(b), _fp(c), M, N, K)
    return c


def row_stats64(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    rows, cols = a.shape
    s = np.zeros(rows, np.float64)
    sa = np.zeros(rows, np.float64)
    m = np.zeros(rows, np.float32)
    t = np.zeros(1, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    lib().fp64_row_stats(_fp(a), rows, cols, s.ctypes.data_as(dp), _fp(m), sa.ctypes.data_as(dp),
                         t.ctypes.data_as(dp))
    return s, m, sa, float(t[0])
