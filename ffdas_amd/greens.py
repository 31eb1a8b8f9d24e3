"""Green's function summation: ``greens`` with the reference's Python signature and errors
(bindings/python/ffdas/greens.py:28-87), on ffdas_greens (include/ffdas.h)."""
from __future__ import annotations

import math

import torch

from ._core import as_tensor, check_device_tensor, element_code, handle_for
from ._lib import TensorDesc, check, lib

__all__ = ["greens"]


def greens(srcpos, wavenums, x, dstpos, *, out=None):
    """Sum over sources of x[..., s, f] exp(i k_f r) / r at every target.

    Args:
        srcpos: source positions, shape (sources, 3).
        wavenums: wavenumber per frequency bin, shape (frequencies,).
        x: complex input field ([batch,] sources, frequencies), complex64 or complex32.
        dstpos: target positions, shape (..., 3).
        out: pre-allocated complex64 output.

    Returns:
        ([batch,] ..., frequencies) complex64, ... the spatial dimensions of dstpos.
    """
    x = as_tensor(x)
    srcpos = as_tensor(srcpos).to(torch.float32).contiguous()
    dstpos = as_tensor(dstpos).to(torch.float32).contiguous()
    wavenums = as_tensor(wavenums).to(torch.float32).contiguous()
    out_shape = (*dstpos.shape[:-1], x.shape[-1])
    flat_shape = (math.prod(dstpos.shape[:-1]), x.shape[-1])
    if x.ndim == 3:
        out_shape = (x.shape[0],) + tuple(out_shape)
        flat_shape = (x.shape[0],) + flat_shape
    elif x.ndim == 2:
        flat_shape = (1,) + flat_shape
        x = x.reshape((1,) + tuple(x.shape))
    else:
        raise ValueError(f"input must have 2 or 3 dimensions, got shape {tuple(x.shape)}")
    out_shape = tuple(int(d) for d in out_shape)
    for name, t in (("x", x), ("srcpos", srcpos), ("dstpos", dstpos), ("wavenums", wavenums)):
        check_device_tensor(t, name, x.device if name != "x" else None)
    if out is None:
        out = torch.empty(out_shape, dtype=torch.complex64, device=x.device)
    elif tuple(out.shape) != out_shape:
        raise ValueError(f"invalid output shape: expected {out_shape}, got {tuple(out.shape)}")
    else:
        check_device_tensor(out, "out", x.device)
    xc = x.contiguous()
    out3 = out.view(flat_shape) if out.is_contiguous() else None
    tmp = out3 if out3 is not None else torch.empty(flat_shape, dtype=out.dtype, device=out.device)
    xd = TensorDesc(tuple(xc.shape), tuple(xc.stride()), element_code(xc))
    od = TensorDesc(tuple(tmp.shape), tuple(tmp.stride()), element_code(tmp))
    with torch.cuda.device(x.device):
        h = handle_for(xc)
        check(lib().ffdas_greens(h.h, srcpos.data_ptr(), wavenums.data_ptr(), xd.ptr,
                                 xc.data_ptr(), dstpos.data_ptr(), od.ptr, tmp.data_ptr()),
              "ffdas_greens")
    if out3 is None:
        out.copy_(tmp.view(out_shape))
    return out
