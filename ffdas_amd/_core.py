"""Handles, tensor views and dtype mapping between torch/DLPack tensors and the C-ABI.

Mirrors the reference's binding layer: one library handle per device
(bindings/python/ffdas/_core/library.py:17-32), a ``TensorView`` that reinterprets an array
with a trailing dimension of 2 as complex half/bfloat16 elements
(bindings/python/ffdas/_core/src/_ffdas.cpp:102-109, 276-306), and DLPack import of device
arrays (the reference imports CUDA arrays; here ROCm device arrays, kDLROCM).
"""
from __future__ import annotations

import atexit
import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import DType, check, lib


# --------------------------------------------------------------------------------------
# element types
# --------------------------------------------------------------------------------------
@dataclass(frozen=True)
class ElementType:
    """A library element type usable with :func:`view`."""
    name: str
    code: int
    base: torch.dtype          # dtype of the underlying torch tensor
    pair: bool                 # underlying tensor carries a trailing dimension of 2

    def __repr__(self):
        return f"ffdas_amd.{self.name}"


half = ElementType("half", DType.R_16F, torch.float16, False)
bfloat16 = ElementType("bfloat16", DType.R_16BF, torch.bfloat16, False)
float32 = ElementType("float32", DType.R_32F, torch.float32, False)
float64 = ElementType("float64", DType.R_64F, torch.float64, False)
half2 = ElementType("half2", DType.C_16F, torch.float16, True)
bfloat162 = ElementType("bfloat162", DType.C_16BF, torch.bfloat16, True)
float2 = ElementType("float2", DType.C_32F, torch.float32, True)
double2 = ElementType("double2", DType.C_64F, torch.float64, True)

_TORCH_CODE = {
    torch.float16: DType.R_16F,
    torch.bfloat16: DType.R_16BF,
    torch.float32: DType.R_32F,
    torch.float64: DType.R_64F,
    torch.complex32: DType.C_16F,
    torch.complex64: DType.C_32F,
    torch.complex128: DType.C_64F,
}


class TensorView:
    """A device tensor reinterpreted as a library element type (reference _ffdas.cpp:276-306).

    For complex pair types the underlying tensor has a trailing dimension of 2 (re, im) which
    is not part of ``shape``.
    """

    def __init__(self, base: torch.Tensor, dtype: ElementType):
        if base.dtype != dtype.base:
            raise ValueError(f"view: expected base element type compatible with {dtype.name}")
        if dtype.pair:
            if base.ndim == 0 or base.shape[-1] != 2:
                raise ValueError("view: complex dtype requires trailing dimension of 2")
            if base.stride(-1) != 1:
                raise ValueError("view: the (re, im) pair must be contiguous")
            self.shape = tuple(base.shape[:-1])
            self.strides = tuple(s // 2 for s in base.stride()[:-1])
            if any(s % 2 for s in base.stride()[:-1]):
                raise ValueError("view: strides must be whole complex elements")
        else:
            self.shape = tuple(base.shape)
            self.strides = tuple(base.stride())
        self.base = base
        self.dtype = dtype

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def device(self):
        return self.base.device

    def data_ptr(self):
        return self.base.data_ptr()

    def __repr__(self):
        return f"TensorView(shape={self.shape}, dtype={self.dtype})"


def view(a, dtype: ElementType) -> TensorView:
    """Reinterpret ``a`` (a torch tensor or any DLPack device array) as ``dtype``."""
    return TensorView(as_tensor(a), dtype)


def as_tensor(a):
    """Import a torch tensor or any ``__dlpack__`` device array without copying."""
    if isinstance(a, (torch.Tensor, TensorView)):
        return a
    if hasattr(a, "__dlpack__"):
        return torch.from_dlpack(a)
    raise TypeError(f"expected a device tensor or DLPack array, got {type(a).__name__}")


def element_code(t) -> int:
    if isinstance(t, TensorView):
        return int(t.dtype.code)
    code = _TORCH_CODE.get(t.dtype)
    if code is None:
        raise TypeError(f"unsupported element type {t.dtype}")
    return int(code)


def shape_strides(t):
    if isinstance(t, TensorView):
        return t.shape, t.strides
    return tuple(t.shape), tuple(t.stride())


def check_device_tensor(t, name: str, device=None):
    dev = t.device
    if dev.type != "cuda":
        raise TypeError(f"{name} must be a ROCm device tensor (got device {dev})")
    if device is not None and dev != device:
        raise ValueError(f"{name} is on {dev}, expected {device}")


# --------------------------------------------------------------------------------------
# handles
# --------------------------------------------------------------------------------------
class Handle:
    """Owns one ffdas_handle_t (reference _ffdas.cpp:118-137)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        check(lib().ffdas_create(ctypes.byref(self.h)), "ffdas_create")

    def destroy(self):
        if self.h:
            lib().ffdas_destroy(self.h)
            self.h = ctypes.c_void_p()

    def set_stream(self, stream: int):
        check(lib().ffdas_set_stream(self.h, int(stream)), "ffdas_set_stream")

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


_handles: dict[int, Handle | None] = {}


def device_get() -> int:
    d = ctypes.c_int()
    check(lib().ffdas_device_get(ctypes.byref(d)), "ffdas_device_get")
    return d.value


def device_set(device: int) -> None:
    check(lib().ffdas_device_set(int(device)), "ffdas_device_set")


def device_count() -> int:
    c = ctypes.c_int()
    check(lib().ffdas_device_count(ctypes.byref(c)), "ffdas_device_count")
    return c.value


def get_library_handle(device: int | None = None) -> Handle:
    """One handle per device, created on first use (reference library.py:17-23)."""
    if device is None:
        device = torch.cuda.current_device()
    h = _handles.get(device)
    if h is None:
        with torch.cuda.device(device):
            h = Handle()
        _handles[device] = h
    return h


def handle_for(t) -> Handle:
    """Handle of t's device, bound to torch's current stream on that device."""
    dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
    h = get_library_handle(dev)
    h.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return h


@atexit.register
def _cleanup() -> None:
    for d in list(_handles):
        h = _handles.get(d)
        if h is not None:
            h.destroy()
        _handles[d] = None
