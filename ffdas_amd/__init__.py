"""ffdas for MI355X: delay-and-sum (``das``, ``das_sparse``) on hand-written gfx950 kernels.

The operators keep the reference's Python signature (bindings/python/ffdas/das.py) and call
the C-ABI library libffdas.so (include/ffdas.h) built from ffdas_amd/csrc.
"""
from ._core import (TensorView, bfloat16, bfloat162, complex64, complex128, device_count,
                    device_get, device_set, double2, float2, float32, float64,
                    get_library_handle, half, half2, int2, int16, int32, short2, view)
from ._lib import Algorithm, ComputeType, DType, FfdasError, InterpMode
from .autograd import DasFunction, das_autograd
from .das import (CapturedCall, DowncastWarning, autotune, clear_tuning, das, das_backward,
                  das_batched, das_sparse, load_tuning, save_tuning, tuning_entries)
from .greens import greens
from .interpolation import Interpolator, interpolate
from .rank import truncate_rank
from .tensor import contiguous_copy
from .utils.spatial import angle, cdist, rect_dist

__version__ = "0.1.0"

__all__ = [
    "das", "das_sparse", "das_backward", "das_autograd", "DasFunction", "Algorithm",
    "ComputeType", "DType", "FfdasError", "DowncastWarning", "TensorView", "view", "half",
    "half2", "bfloat16", "bfloat162", "float32", "float64", "float2", "double2",
    "get_library_handle", "device_count", "device_get", "device_set", "autotune",
    "save_tuning", "load_tuning", "clear_tuning", "tuning_entries", "das_batched",
    "CapturedCall", "cdist", "rect_dist", "angle", "truncate_rank", "Interpolator",
    "interpolate", "greens", "int16", "int32", "short2", "int2", "complex64", "complex128",
    "contiguous_copy", "InterpMode",
]
