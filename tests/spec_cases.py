"""The reference's known-answer cases (TensorsSpec, cpuSpec, cpu.scala and README examples)
written once against the Tensor API, so the same case runs through the oracle on the CPU
and through the gfx950 back-end on the GPU. Each case returns a value comparable with the
`expect` of tests/golden/known_answers.json."""
import json
import os
from functools import reduce

import numpy as np

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def convolute(T, inp, weight, bias):
    """TensorsSpec.scala:144-210."""
    batch, height, width, depth = inp.shape
    kh, kw, depth2, filters = weight.shape
    assert depth2 == depth and bias.shape == (filters,)
    input_seq = inp.split(3)
    weight_seq = [[[list(d.split(0)) for d in kwd.split(0)] for kwd in khkwd.split(0)] for khkwd in weight.split(3)]
    bias_seq = bias.split(0)
    channels = []
    for weight_per_filter, bias_per_filter in zip(weight_seq, bias_seq):
        summands = []
        for offset_y, weight_per_row in zip(range(-1, 2), weight_per_filter):
            for offset_x, weight_per_pixel in zip(range(-1, 2), weight_per_row):
                for input_per_channel, weight_per_channel in zip(input_seq, weight_per_pixel):
                    assert weight_per_channel.shape == ()
                    summands.append(input_per_channel.translate([0, offset_y, offset_x]) *
                                    weight_per_channel.broadcast((batch, height, width)))
        channels.append(bias_per_filter.broadcast((batch, height, width)) + reduce(lambda a, b: a + b, summands))
    return T.Tensor.join(channels)


def mm_broadcast(T, a, b):
    """matrixMultiply of TensorsSpec.scala:472-479."""
    i, j = a.shape
    j2, k = b.shape
    product = a.broadcast((i, j, k)) * b.reshape((1, j, k)).broadcast((i, j, k))
    return reduce(lambda x, y: x + y, list(product.split(1)))


def mm_unrolled(T, a, b):
    """matrixMultiply of TensorsSpec.scala:506-518."""
    columns1 = a.split(1)
    result_columns = []
    for column2 in b.split(1):
        terms = [l * r.broadcast(l.shape) for l, r in zip(columns1, column2.split(0))]
        result_columns.append(reduce(lambda x, y: x + y, terms))
    return T.Tensor.join(result_columns)


M1 = [[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]
M2 = [[7.0, 8.0, 9.0, 10.0], [11.0, 12.0, 13.0, 14.0], [15.0, 16.0, 17.0, 18.0]]


def build(name, T):
    """Returns (tensor or list of tensors, how) for a known-answer case."""
    Tensor = T.Tensor
    if name == "repeated_literal":
        return Tensor(42.0), "string"
    if name == "literal_1d":
        return Tensor([1.0, 2.0]), "string"
    if name == "literal_2d":
        return Tensor([[1.0, 2.0], (3.0, 4.0)]), "string"
    if name == "translate_fill":
        return Tensor.fill(42.0, (2, 3, 5), padding=99.0).translate([1, 2, -3]), "string"
    if name == "unzip":
        return list(Tensor([[[[1.0, 5.0]]]]).split(3)), "strings"
    if name == "plus":
        t = Tensor([[[1.0, 5.0]]])
        return t + t, "string"
    if name == "plus_times":
        t = Tensor([[[1.0, 5.0]]])
        t2 = t + t
        return t2 * t2, "string"
    if name == "convolution":
        inp = np.zeros((2, 4, 5, 3), np.float32)
        inp[0, 0, 0, 0] = 1.0
        inp[0, 1, 0, 0] = 10.0
        inp[1, 0, 0, 0] = 100.0
        w = np.zeros((3, 3, 3, 2), np.float32)
        w[1, 1, 0, 0] = 3.0
        w[1, 1, 0, 1] = 4.0
        w[0, 1, 0, 0] = 5.0
        w[2, 2, 0, 1] = 6.0
        out = convolute(T, Tensor(inp), Tensor(w), Tensor(np.array([100000.0, 200000.0], np.float32)))
        assert out.shape == (2, 4, 5, 2)
        return out, "points"
    if name == "sum_fill":
        return Tensor.fill(15625.0, (8, 8)).sum(), "string"
    if name == "random_normal_scalar":
        return Tensor.randomNormal((), seed=54321), "flat"
    if name == "read_scalar":
        return Tensor(42.0), "scalar"
    if name == "read_1d":
        return Tensor([1.0, 2.0]), "nested"
    if name == "read_2d":
        return Tensor([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]]), "nested"
    if name == "read_3d":
        return Tensor(GOLDEN["read_3d"]["expect"]), "nested"
    if name == "read_4d":
        return Tensor(GOLDEN["read_4d"]["expect"]), "nested"
    if name == "random":
        return Tensor.random((3, 3), seed=12345), "string"
    if name == "random_normal":
        return Tensor.randomNormal((3, 3, 3), seed=54321), "string"
    if name == "transpose_0d":
        return Tensor(42.0).transpose(), "string"
    if name == "transpose_1d":
        return Tensor([1.0, 2.0, 3.0]).transpose(), "string"
    if name == "transpose_2d":
        return Tensor([[1.0, 2.0], [3.0, 4.0]]).transpose(), "string"
    if name == "transpose_3d":
        return Tensor([[[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], [[7.0, 8.0, 9.0], [10.0, 11.0, 12.0]]]).transpose(), "string"
    if name == "matrix_multiplication":
        return mm_broadcast(T, Tensor(M1), Tensor(M2)), "string"
    if name == "broadcast":
        return Tensor(M1).broadcast((2, 3, 4)), "string"
    if name == "unrolled_matrix_multiplication":
        return mm_unrolled(T, Tensor(M1), Tensor(M2)), "string"
    if name == "computational_graph":
        a = Tensor.fill(2.0, (2, 3)).nonInline
        b = Tensor.fill(2.0, (2, 3)).nonInline
        c = (a + b).nonInline
        d = (c + b).nonInline
        return d, "string"
    if name.startswith("join_"):
        a = Tensor.fill(42.0, (3, 4))
        b = Tensor.fill(43.0, (3, 4))
        return Tensor.join([a, b], int(name[-1])), "string+shape"
    if name == "doc_2d_literal":
        return Tensor([(1.0, 2.0), (3.0, 4.0)]), "string"
    if name in ("doc_split_0", "doc_split_1"):
        t = Tensor(np.arange(24, dtype=np.float32).reshape(2, 3, 4).tolist())
        assert t.shape == (2, 3, 4)
        return t.split(int(name[-1])), "tensorseq"
    if name == "doc_join":
        parts = [Tensor([[1.0, 2.0], [3.0, 4.0]]), Tensor([[5.0, 6.0], [7.0, 8.0]]),
                 Tensor([[9.0, 10.0], [11.0, 12.0]])]
        return Tensor.join(parts), "string+shape"
    if name == "doc_scalar_broadcast":
        return Tensor.scalar(42.0).broadcast((2, 3)), "string"
    if name == "readme_2d":
        return Tensor([(1.0, 2.0, 3.0), (4.0, 5.0, 6.0)]), "string"
    if name == "readme_plus100":
        m = Tensor([(1.0, 2.0, 3.0), (4.0, 5.0, 6.0)])
        return m + Tensor.fill(100.0, (2, 3)), "string"
    raise KeyError(name)


CASES = [k for k in GOLDEN if not k.startswith("_")]


def check(name, T, to_string, flat):
    """Builds case `name` on `T`, renders it with the given functions and compares with the
    golden value. Returns nothing; raises AssertionError on mismatch."""
    g = GOLDEN[name]
    obj, how = build(name, T)
    exp = g["expect"]
    if how == "string":
        assert to_string(obj) == exp, (name, to_string(obj), exp)
    elif how == "string+shape":
        assert list(obj.shape) == g["shape"]
        assert to_string(obj) == exp
    elif how == "strings":
        assert [to_string(t) for t in obj] == exp
    elif how == "tensorseq":
        assert "TensorSeq(" + ", ".join(to_string(t) for t in obj) + ")" == exp
    elif how == "points":
        a = flat(obj).reshape(obj.shape)
        for idx, v in exp:
            assert a[tuple(idx)] == np.float32(v), (idx, a[tuple(idx)], v)
    elif how == "flat":
        assert flat(obj).tolist() == [float(np.float32(v)) for v in exp]
    elif how == "scalar":
        assert float(flat(obj)[0]) == exp
    elif how == "nested":
        assert flat(obj).reshape(obj.shape).tolist() == exp
    else:
        raise KeyError(how)
