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
    """A library element type usable with :func:`view` (reference _ffdas.cpp:57-70)."""
    name: str
    code: int
    base: torch.dtype          # dtype of the underlying torch tensor
    pair: bool                 # underlying tensor carries a trailing dimension of 2
    dl_code: int = 2           # DLPack type code of one element (0 int, 2 float, 4 bfloat,
    dl_bits: int = 32          # 5 complex), bits per lane and lanes
    dl_lanes: int = 1

    def __repr__(self):
        return f"ffdas_amd.{self.name}"


_DL_INT, _DL_FLOAT, _DL_BFLOAT, _DL_COMPLEX = 0, 2, 4, 5

half = ElementType("half", DType.R_16F, torch.float16, False, _DL_FLOAT, 16)
bfloat16 = ElementType("bfloat16", DType.R_16BF, torch.bfloat16, False, _DL_BFLOAT, 16)
float32 = ElementType("float32", DType.R_32F, torch.float32, False, _DL_FLOAT, 32)
float64 = ElementType("float64", DType.R_64F, torch.float64, False, _DL_FLOAT, 64)
int16 = ElementType("int16", DType.R_16I, torch.int16, False, _DL_INT, 16)
int32 = ElementType("int32", DType.R_32I, torch.int32, False, _DL_INT, 32)
half2 = ElementType("half2", DType.C_16F, torch.float16, True, _DL_FLOAT, 16, 2)
bfloat162 = ElementType("bfloat162", DType.C_16BF, torch.bfloat16, True, _DL_BFLOAT, 16, 2)
float2 = ElementType("float2", DType.C_32F, torch.float32, True, _DL_FLOAT, 32, 2)
double2 = ElementType("double2", DType.C_64F, torch.float64, True, _DL_FLOAT, 64, 2)
short2 = ElementType("short2", DType.C_16I, torch.int16, True, _DL_INT, 16, 2)
int2 = ElementType("int2", DType.C_32I, torch.int32, True, _DL_INT, 32, 2)
complex64 = ElementType("complex64", DType.C_32F, torch.complex64, False, _DL_COMPLEX, 64)
complex128 = ElementType("complex128", DType.C_64F, torch.complex128, False, _DL_COMPLEX, 128)

_TORCH_CODE = {
    torch.float16: DType.R_16F,
    torch.bfloat16: DType.R_16BF,
    torch.float32: DType.R_32F,
    torch.float64: DType.R_64F,
    torch.int16: DType.R_16I,
    torch.int32: DType.R_32I,
    torch.complex32: DType.C_16F,
    torch.complex64: DType.C_32F,
    torch.complex128: DType.C_64F,
}


# --------------------------------------------------------------------------------------
# DLPack export (the legacy "dltensor" capsule, DLPack 0.x ABI)
# --------------------------------------------------------------------------------------
class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DL_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", _DL_DELETER)]

KDL_CPU, KDL_ROCM = 1, 10
# exported tensors alive until their consumer (or the unconsumed capsule) calls the deleter:
# address of the DLManagedTensor -> (struct, shape array, strides array, base tensor)
_dl_live: dict[int, tuple] = {}


@_DL_DELETER
def _dl_deleter(mt):
    _dl_live.pop(ctypes.addressof(mt.contents), None)


_CAPSULE_DTOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
# private prototypes: the attributes of ctypes.pythonapi.* are shared with other code
_capsule_new = ctypes.PYFUNCTYPE(ctypes.py_object, ctypes.c_void_p, ctypes.c_char_p,
                                 _CAPSULE_DTOR)(("PyCapsule_New", ctypes.pythonapi))
_capsule_valid = ctypes.PYFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p)(
    ("PyCapsule_IsValid", ctypes.pythonapi))
_capsule_ptr = ctypes.PYFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p)(
    ("PyCapsule_GetPointer", ctypes.pythonapi))


@_CAPSULE_DTOR
def _capsule_dtor(cap):
    # an unconsumed capsule keeps its name "dltensor": release the export; a consumer renames
    # it ("used_dltensor") and calls the deleter itself
    if _capsule_valid(cap, b"dltensor"):
        addr = _capsule_ptr(cap, b"dltensor")
        if addr:
            _dl_live.pop(addr, None)


def dlpack_device_of(t: torch.Tensor) -> tuple[int, int]:
    if t.device.type == "cuda":
        return KDL_ROCM, t.device.index if t.device.index is not None else 0
    return KDL_CPU, 0


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
        shp = "(" + ", ".join(map(str, self.shape)) + ("," if len(self.shape) == 1 else "") + ")"
        return f"TensorView(shape={shp}, dtype={self.dtype!r})"

    def __dlpack_device__(self) -> tuple[int, int]:
        """(device type, device id): kDLROCM (10) for a ROCm tensor (the reference reports
        kDLCUDA for its CUDA arrays, _ffdas.cpp:229-231)."""
        return dlpack_device_of(self.base)

    def __dlpack__(self, stream=None, **kwargs):
        """Export the view as a DLPack capsule with the view's element type (reference
        _ffdas.cpp:232-275): complex pair types as two-lane vectors (half2: float16 x 2),
        complex64/complex128 as DLPack complex, the view's shape and its element strides.  The
        base tensor stays alive until the consumer releases the capsule.  The data is not
        copied or synchronised: `stream` is accepted and ignored, as work on the base tensor
        is already ordered on torch's current stream."""
        nd = len(self.shape)
        shape = (ctypes.c_int64 * max(nd, 1))(*self.shape)
        strides = (ctypes.c_int64 * max(nd, 1))(*self.strides)
        mt = _DLManagedTensor()
        mt.dl_tensor.data = self.base.data_ptr()
        dt, di = self.__dlpack_device__()
        mt.dl_tensor.device = _DLDevice(dt, di)
        mt.dl_tensor.ndim = nd
        mt.dl_tensor.dtype = _DLDataType(self.dtype.dl_code, self.dtype.dl_bits,
                                         self.dtype.dl_lanes)
        mt.dl_tensor.shape = ctypes.cast(shape, ctypes.POINTER(ctypes.c_int64))
        mt.dl_tensor.strides = ctypes.cast(strides, ctypes.POINTER(ctypes.c_int64))
        mt.dl_tensor.byte_offset = 0
        mt.manager_ctx = None
        mt.deleter = _dl_deleter
        addr = ctypes.addressof(mt)
        _dl_live[addr] = (mt, shape, strides, self.base)
        return _capsule_new(addr, b"dltensor", _capsule_dtor)


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
