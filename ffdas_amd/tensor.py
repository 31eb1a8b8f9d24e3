"""``contiguous_copy``: the reference's strided copy with type conversion
(bindings/python/ffdas/_core/tensor.py:174-191) over ffdas_contiguous_copy."""
from __future__ import annotations

import ctypes

import torch

from . import _core
from ._core import TensorView, as_tensor, check_device_tensor, element_code, handle_for
from ._lib import TensorDesc, check, lib


def _empty_like(x):
    if isinstance(x, TensorView):
        shape = tuple(x.shape) + ((2,) if x.dtype.pair else ())
        return TensorView(torch.empty(shape, dtype=x.dtype.base, device=x.device), x.dtype)
    return torch.empty(tuple(x.shape), dtype=x.dtype, device=x.device)


def contiguous_copy(x, out=None):
    """Copy a strided device array into a contiguous one.

    Args:
        x: input on a ROCm device (torch tensor, :class:`TensorView` or DLPack array; any
            strides, negative included).
        out: optional preallocated contiguous output with x's shape; its element type selects
            the conversion (int16 -> float32, float32 <-> float16 / float64, float64 ->
            float32, complex likewise, plus bfloat16).

    Returns:
        ``out`` (a new contiguous array of x's type when not given).
    """
    x = as_tensor(x)
    check_device_tensor(x, "x")
    if out is None:
        out = _empty_like(x)
    else:
        out = as_tensor(out)
        check_device_tensor(out, "out", x.device)
    xs, xst = _core.shape_strides(x)
    os_, ost = _core.shape_strides(out)
    xd = TensorDesc(xs, xst, element_code(x))
    od = TensorDesc(os_, ost, element_code(out))
    with torch.cuda.device(x.device):
        h = handle_for(x)
        check(lib().ffdas_contiguous_copy(h.h, xd.ptr, ctypes.c_void_p(x.data_ptr()), od.ptr,
                                          ctypes.c_void_p(out.data_ptr())),
              "ffdas_contiguous_copy")
    return out
