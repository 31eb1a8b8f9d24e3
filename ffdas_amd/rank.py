"""Rank truncation filter: ``truncate_rank`` with the reference's Python signature and errors
(bindings/python/ffdas/truncate_rank.py:28-82), on ffdas_truncate_rank (include/ffdas.h)."""
from __future__ import annotations

import math

import torch

from ._core import as_tensor, check_device_tensor, element_code, handle_for
from ._lib import TensorDesc, check, lib

__all__ = ["truncate_rank"]


def truncate_rank(x, start: int, stop: int | None = None, *, out=None):
    """Rank truncation filter via truncated SVD reconstruction.

    Reconstructs x from its singular vectors start through stop - 1 (0-based, counted from
    the largest singular value), removing all other components.  x (at least 2-D; float32,
    complex64, float64 or complex128 on the GPU) is treated as the matrix
    (x.shape[0], prod(x.shape[1:])) in C order; the result has x's shape and dtype.

    Args:
        x: input array, at least 2 dimensions.
        start: index of the first singular vector to keep (0-based).
        stop: index past the last singular vector to keep; defaults to min(m, n).
        out: pre-allocated output of x's shape (returned).
    """
    x = as_tensor(x)
    if x.ndim < 2:
        raise ValueError(f"x must have at least 2 dimensions, got {x.ndim}")
    m = int(x.shape[0])
    n = math.prod(int(d) for d in x.shape[1:])
    if start < 0 or start >= min(m, n):
        raise ValueError(f"start must be >= 0 and < min(m, n) ({m}, {n}), got {start}")
    if stop is None:
        stop = min(m, n)
    elif stop <= start or stop > min(m, n):
        raise ValueError(
            f"stop must be > start ({start}) and <= min(m, n) ({m}, {n}), got {stop}")
    if out is not None and tuple(out.shape) != tuple(x.shape):
        raise ValueError("out must have the same dimensions as x")

    check_device_tensor(x, "x")
    result = out
    if out is None:
        out = torch.empty(x.shape, dtype=x.dtype, device=x.device)
        result = out
    else:
        check_device_tensor(out, "out", x.device)
    x2 = x.reshape(m, n)
    try:
        out2 = out.view(m, n)
        tmp = None
    except RuntimeError:              # an out that cannot be seen as (m, n): write, then copy
        tmp = torch.empty((m, n), dtype=out.dtype, device=out.device)
        out2 = tmp
    x_desc = TensorDesc((m, n), tuple(x2.stride()), element_code(x2))
    o_desc = TensorDesc((m, n), tuple(out2.stride()), element_code(out2))
    with torch.cuda.device(x.device):
        h = handle_for(x)
        check(lib().ffdas_truncate_rank(h.h, x_desc.ptr, x2.data_ptr(), int(start), int(stop),
                                        o_desc.ptr, out2.data_ptr()), "ffdas_truncate_rank")
    if tmp is not None:
        result.copy_(tmp.view(result.shape))
    return result
