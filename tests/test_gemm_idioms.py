"""The reference's matrix-multiply idioms (README.md:301-362, TensorsSpec.scala:468-528) are
recognised from the closures they build, without a device (CPU)."""
import numpy as np
import pytest

from compute import Tensors, tensors as tz, trees


def mm1(T, m1, m2):
    """matrixMultiply1 as the reference writes it."""
    import functools
    columns1 = list(m1.split(1))
    columns2 = list(m2.split(1))
    result_columns = []
    for column2 in columns2:
        prods = [l * r.broadcast(l.shape) for l, r in zip(columns1, column2.split(0))]
        result_columns.append(functools.reduce(lambda a, b: a + b, prods))
    return T.Tensor.join(result_columns)


def mm2(T, m1, m2):
    """matrixMultiply2 as the reference writes it."""
    import functools
    i, j = m1.shape
    j2, k = m2.shape
    assert j == j2
    product = m1.broadcast((i, j, k)) * m2.reshape((1, j, k)).broadcast((i, j, k))
    return functools.reduce(lambda a, b: a + b, list(product.split(1)))


@pytest.fixture(scope="module")
def T():
    return Tensors(offline=True)


@pytest.mark.parametrize("shape", [(2, 3, 4), (64, 8, 8), (5, 2, 7)])
def test_mm2_is_recognised(T, shape):
    i, j, k = shape
    a = T.Tensor(np.zeros((i, j), np.float32))
    b = T.Tensor(np.zeros((j, k), np.float32))
    c = mm2(T, a, b)
    fold = tz._split_fold(c.closure, c.shape)
    assert fold is not None and fold[1] == 1 and fold[2] == "+"
    got = tz._match_mm2(fold[0].id)
    assert got is not None and got[0] is a and got[2:] == (i, j, k)
    assert tuple(got[1].shape) == (1, j, k) and got[1]._source is b


@pytest.mark.parametrize("shape", [(2, 3, 4), (64, 8, 8), (5, 1, 7)])
def test_mm1_is_recognised(T, shape):
    i, j, k = shape
    a = T.Tensor(np.zeros((i, j), np.float32))
    b = T.Tensor(np.zeros((j, k), np.float32))
    c = mm1(T, a, b)
    got = tz._match_mm1(trees.Concatenate([p.closure for p in c._parts]), c._head_shape)
    assert got is not None and got[0] is a and got[1] is b and got[2:] == (i, j, k)


def test_near_misses_are_not_recognised(T):
    """Operands swapped inside the fold, a translated column, a max fold, a different
    broadcast: none is a matrix product, none is matched."""
    import functools
    a = T.Tensor(np.zeros((4, 3), np.float32))
    b = T.Tensor(np.zeros((3, 5), np.float32))
    # a translated slice of A in one term
    cols1 = list(a.split(1))
    cols2 = list(b.split(1))
    parts = []
    for col2 in cols2:
        prods = [l * r.broadcast(l.shape) for l, r in zip(cols1, col2.split(0))]
        prods[1] = cols1[1].translate([1]) * col2.split(0)[1].broadcast((4,))
        parts.append(functools.reduce(lambda x, y: x + y, prods))
    assert tz._match_mm1(trees.Concatenate([p.closure for p in parts]), (4,)) is None
    # B's rows out of order
    parts = []
    for col2 in cols2:
        rs = list(col2.split(0))
        prods = [l * r.broadcast(l.shape) for l, r in zip(cols1, reversed(rs))]
        parts.append(functools.reduce(lambda x, y: x + y, prods))
    assert tz._match_mm1(trees.Concatenate([p.closure for p in parts]), (4,)) is None
    # a max fold over the product, and a sum of the product plus a constant
    product = a.broadcast((4, 3, 5)) * b.reshape((1, 3, 5)).broadcast((4, 3, 5))
    f = functools.reduce(T.Tensor.max, list(product.split(1)))
    fold = tz._split_fold(f.closure, f.shape)
    assert fold is not None and fold[2] == "max"
    shifted = product + T.Tensor.fill(1.0, (4, 3, 5))
    g = functools.reduce(lambda x, y: x + y, list(shifted.split(1)))
    fold = tz._split_fold(g.closure, g.shape)
    assert fold is not None and tz._match_mm2(fold[0].id) is None
