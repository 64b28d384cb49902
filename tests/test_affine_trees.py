"""NDimensionalAffineTransform and Trees mirrors (CPU).

Affine composition is pinned by the 2-D closed form (java.awt.geom.AffineTransform
semantics of NDimensionalAffineTransformSpec.scala:14-56: a 2x3 matrix is the top of a
3x3 homogeneous matrix); structural equality by TreesSpec.scala:13-93's cases."""
import numpy as np
import pytest

from compute import affine, trees


def homogeneous(m, rows, cols_in):
    h = np.zeros((rows + 1, cols_in + 1))
    h[:rows] = np.reshape(m, (rows, cols_in + 1))
    h[rows, cols_in] = 1
    return h


@pytest.mark.parametrize("seed", range(5))
def test_pre_concatenate_2d_closed_form(seed):
    rng = np.random.default_rng(seed)
    m0 = list(rng.integers(0, 100, 6).astype(float)) if seed else [1.0, 0.0, 3.5, 0.0, 1.0, 4.2]
    m1 = list(rng.integers(0, 100, 6).astype(float)) if seed else [3.0, 0.0, 0.0, 0.0, 2.0, 0.0]
    got = affine.pre_concatenate(m0, m1, 2)  # m1 after m0
    want = (homogeneous(m1, 2, 2) @ homogeneous(m0, 2, 2))[:2].ravel()
    assert np.allclose(got, want)
    got2 = affine.concatenate(m1, m0, 2)
    assert np.allclose(got2, want)


def test_non_square_composition_and_transform():
    # 3-D -> 2-D view, then a 2-D -> 4-D view
    m01 = [1, 0, 0, 5, 0, 0, 1, -1]  # (2 rows, 4 cols): space0 is 3-D
    m12 = list(np.arange(12, dtype=float))  # (4 rows, 3 cols): space1 is 2-D
    got = affine.pre_concatenate(m01, m12, 3)
    want = (homogeneous(m12, 4, 2) @ homogeneous(m01, 2, 3))[:4].ravel()
    assert np.allclose(got, want)
    p = [2.0, 7.0, -1.0]
    assert np.allclose(affine.transform(got, p), affine.transform(m12, affine.transform(m01, p)))
    assert affine.translate([1, -2]) == [1.0, 0.0, 1.0, 0.0, 1.0, -2.0]


def leaf(ident, shape=(2, 3), padding=0.0):
    return trees.ArrayParameter(ident, padding, shape)


def test_structural_equality_and_alpha_conversion():
    a, b, c = leaf("a"), leaf("b"), leaf("c")
    e1 = trees.Binary("+", trees.Extract(a), trees.Extract(b))
    e2 = trees.Binary("+", trees.Extract(c), trees.Extract(a))  # same structure, renamed
    assert trees.is_same_structure(e1, e2)
    assert trees.structural_hash(e1) == trees.structural_hash(e2)
    e3 = trees.Binary("+", trees.Extract(a), trees.Extract(a))  # one leaf used twice
    assert not trees.is_same_structure(e1, e3)
    e4 = trees.Binary("*", trees.Extract(a), trees.Extract(b))
    assert not trees.is_same_structure(e1, e4)
    assert not trees.is_same_structure(trees.Extract(leaf("a", (2, 3))), trees.Extract(leaf("a", (3, 2))))
    assert not trees.is_same_structure(trees.Extract(leaf("a", padding=0.0)), trees.Extract(leaf("a", padding=1.0)))
    assert trees.is_same_structure(trees.FloatLiteral(1.5), trees.FloatLiteral(1.5))
    assert not trees.is_same_structure(trees.FloatLiteral(1.5), trees.FloatLiteral(2.5))
    t1 = trees.Extract(trees.Transform(a, [0, 1, 0, 1, 0, 0], 2))
    t2 = trees.Extract(trees.Transform(b, [0, 1, 0, 1, 0, 0], 2))
    t3 = trees.Extract(trees.Transform(b, [0, 1, 1, 1, 0, 0], 2))
    assert trees.is_same_structure(t1, t2) and not trees.is_same_structure(t1, t3)
    j1 = trees.Concatenate([trees.Extract(a), trees.Extract(b)])
    j2 = trees.Concatenate([trees.Extract(b), trees.Extract(c)])
    assert trees.is_same_structure(j1, j2)


def test_parameter_order_is_first_visit():
    a, b, c = leaf("a"), leaf("b"), leaf("c")
    e = trees.Binary("+", trees.Binary("*", trees.Extract(b), trees.Extract(a)),
                     trees.Binary("-", trees.Extract(a), trees.Extract(c)))
    assert [p.id for p in trees.parameters(e)] == ["b", "a", "c"]
    nodes, outputs, leaves = trees.serialize(e)
    assert [p.id for p in leaves] == ["b", "a", "c"]
    slots = [n.param for n in nodes if n.op == 2]
    assert slots == [0, 1, 2]


def test_split_fold_recognised(offline_tensors):
    """split(d) folded with + or max over one materialised tensor is an axis reduction;
    folds out of order, with mixed ops, over views of views or over several tensors are not."""
    import functools
    from compute.tensors import _split_fold
    T = offline_tensors
    t = T.Tensor(np.zeros((3, 5, 4), np.float32))
    for d in range(3):
        for op, f in (("+", lambda a, b: a + b), ("max", T.Tensor.max)):
            r = functools.reduce(f, list(t.split(d)))
            m = _split_fold(r.closure, r.shape)
            assert m is not None and m[1] == d and m[2] == op and m[0].id is t
    s = list(t.split(1))
    assert _split_fold((s[1] + s[0] + s[2] + s[3] + s[4]).closure, (3, 4)) is None  # order
    assert _split_fold(T.Tensor.max(s[0] + s[1] + s[2] + s[3], s[4]).closure, (3, 4)) is None  # mixed
    assert _split_fold((s[0] + s[1] + s[2] + s[3]).closure, (3, 4)) is None  # a slice missing
    v = list(t.transpose().split(0))  # slices of a view: composed matrices
    assert _split_fold(functools.reduce(lambda a, b: a + b, v).closure, (5, 3)) is None
    u = T.Tensor(np.zeros((3, 5, 4), np.float32))
    w = list(u.split(1))
    assert _split_fold((s[0] + s[1] + w[2] + s[3] + s[4]).closure, (3, 4)) is None  # two tensors
