"""Compute.scala's Tensor API on AMD Instinct MI355X (gfx950).

`compute.tensors.Tensors` mirrors `trait Tensors`; `compute.gpu()` is the default
back-end on GPU 0, the counterpart of `import com.thoughtworks.compute.gpu._`.
"""
from . import _capi
from .tensors import (CachedTensor, FillTensor, IllegalArgumentException, InlineTensor, NonInlineTensor,
                      Tensor, Tensors, TensorSeq, TransformedTensor, gpu)

__all__ = ["Tensors", "Tensor", "InlineTensor", "TransformedTensor", "FillTensor", "NonInlineTensor",
           "CachedTensor", "TensorSeq", "IllegalArgumentException", "gpu"]
