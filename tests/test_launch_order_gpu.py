"""The tiled launch order (workgroups take tiles across the output dimension along which a
transposed leaf's rows run, padded where the extents do not divide) computes every element
exactly as the linear order does: the same bits for transposes of rank 2 to 4, several vector
widths, vectors per thread and workgroup sizes, and in a fused expression with other leaves."""
import numpy as np
import pytest

from compute import _capi as capi
from oracle import inputs

pytestmark = pytest.mark.gpu

CASES = [
    ((64, 96), (1, 0)),
    ((4, 64, 96), (2, 0, 1)),
    ((8, 32, 6, 40), (3, 1, 0, 2)),
    ((40, 128), (1, 0)),
    ((5, 12, 64), (2, 0, 1)),
    ((934, 318), (1, 0)),  # padded tiles: 318 and 934 have no large power-of-two factor
    ((9, 7), (1, 0)),
    ((6, 35, 22), (1, 2, 0)),
]
SHAPES = [(4, 1, 256), (2, 2, 512), (1, 4, 128), (8, 1, 1024)]


@pytest.mark.parametrize("shape,perm", CASES)
@pytest.mark.parametrize("V,U,block", SHAPES)
def test_tiled_order_gives_the_linear_bits(shape, perm, V, U, block):
    from compute import Tensors
    x = inputs.uniform(int(np.prod(shape)), 901).reshape(shape)
    out_shape = tuple(shape[p] for p in perm)
    w = inputs.uniform(int(np.prod(out_shape)), 902).reshape(out_shape)
    got, tiled = {}, {}
    for order in (capi.ORDER_AUTO, capi.ORDER_LINEAR):
        with Tensors(0, vector_width=V, items_per_thread=U, block_size=block, launch_order=order) as T:
            xt = T.Tensor(x).permute(list(perm))
            r = T.Tensor.exp(xt) * T.Tensor(w) + xt
            tiled[order] = "dt * " in T.kernel_source(r.closure, r.shape)
            got[order] = r.flatArray()
    print("shape %s perm %s V=%d U=%d block %d: tiled %s" % (shape, perm, V, U, block, tiled[capi.ORDER_AUTO]),
          flush=True)
    assert not tiled[capi.ORDER_LINEAR]
    a, b = got[capi.ORDER_AUTO], got[capi.ORDER_LINEAR]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # and the values are the expression's (exp within a few ulp of the FP64 one)
    xp = np.transpose(x, perm).astype(np.float64)
    ref = np.exp(xp) * w.astype(np.float64) + xp
    assert np.allclose(a.reshape(out_shape), ref, rtol=2e-6, atol=2e-6)


def test_tiling_is_used_where_it_applies():
    from compute import Tensors
    with Tensors(0, vector_width=4) as T:
        r = T.Tensor(np.zeros((64, 96), np.float32)).permute([1, 0]) + T.Tensor.fill(1.0, (96, 64))
        assert "dt * 32u" in T.kernel_source(r.closure, r.shape)
