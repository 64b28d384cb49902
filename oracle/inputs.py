"""Seeded synthetic inputs on the reference's Wang hash (test infrastructure).

The hash is the one `WangHashingRandomNumberGenerator` emits (Tensors.scala:106-117).
  uniform(n, seed, lo, hi):  u = hash(i ^ seed) / 2^32 in FP64, x = lo + (hi - lo) * u,
                             rounded once to FP32
  normal(n, seed):           u1 = (hash + 1) / 2^32, u2 = xorshift(hash) / 2^32,
                             z = sqrt(-2 ln u1) cos(2 pi u2) in FP64, rounded to FP32
These match the library's cs_fill_uniform / cs_fill_normal definitions.
"""
import numpy as np


def wang_hash(v):
    v = np.asarray(v, dtype=np.uint32)
    with np.errstate(over="ignore"):
        v = (v ^ np.uint32(61)) ^ (v >> np.uint32(16))
        v = v * np.uint32(9)
        v = v ^ (v << np.uint32(4))
        v = v * np.uint32(0x27D4EB2D)
        v = v ^ (v >> np.uint32(15))
    return v


def xorshift(s):
    s = np.asarray(s, dtype=np.uint32)
    t1 = s ^ (s << np.uint32(13))
    t2 = t1 ^ (t1 >> np.uint32(17))
    return t2 ^ (t2 << np.uint32(5))


def _index(n, start=0):
    return np.arange(start, start + n, dtype=np.uint64).astype(np.uint32)


def uniform(n, seed, lo=-1.0, hi=1.0, start=0):
    """Elements start .. start + n - 1 of the sequence."""
    h = wang_hash(_index(n, start) ^ np.uint32(seed & 0xFFFFFFFF))
    u = h.astype(np.float64) * (1.0 / 4294967296.0)
    return (lo + (hi - lo) * u).astype(np.float32)


def normal(n, seed):
    h = wang_hash(_index(n) ^ np.uint32(seed & 0xFFFFFFFF))
    u1 = (h.astype(np.float64) + 1.0) * (1.0 / 4294967296.0)
    u2 = xorshift(h).astype(np.float64) * (1.0 / 4294967296.0)
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)).astype(np.float32)


def heavy_tails(x, seed, fraction=0.01, factor=1e4):
    """Entries with hash(i ^ seed) < fraction * 2^32 are multiplied (in FP32) by `factor`."""
    x = np.array(x, dtype=np.float32, copy=True).ravel()
    h = wang_hash(_index(x.size) ^ np.uint32(seed))
    mask = h.astype(np.float64) < fraction * 4294967296.0
    x[mask] = x[mask] * np.float32(factor)
    return x


def row_scales(rows, seed, lo_exp=-3.0, hi_exp=3.0):
    """Per-row scale 10^U(lo_exp, hi_exp) in FP64, rounded to FP32."""
    u = wang_hash(_index(rows) ^ np.uint32(seed)).astype(np.float64) * (1.0 / 4294967296.0)
    return (10.0 ** (lo_exp + (hi_exp - lo_exp) * u)).astype(np.float32)


# The workloads the measurements use, as (name, builder).
def c1_inputs():
    n = 1024 * 1024
    return [uniform(n, s).reshape(1024, 1024) for s in (1001, 1002, 1003)]


def c2_inputs():
    x = normal(64 * 256 * 256, 2001).reshape(64, 256, 256)
    bias = uniform(256, 2002, -0.5, 0.5)
    return x, bias


def c3_input(rows=16384, cols=4096):
    return heavy_tails(normal(rows * cols, 3001), 3002).reshape(rows, cols)


def c4_inputs(n=4096, scaled=False):
    a = uniform(n * n, 4001).reshape(n, n)
    b = uniform(n * n, 4002).reshape(n, n)
    if scaled:
        a = a * row_scales(n, 4003)[:, None]
        b = b * row_scales(n, 4004)[:, None]
    return a.astype(np.float32), b.astype(np.float32)
