"""Element types, enums and TensorView (CPU): the reference's bindings/python/tests/
test_types.py restated (dtype names, reprs, hashing, enum values, views), plus the fields of
the DLPack capsule a view exports (reference _ffdas.cpp:232-275)."""
import ctypes

import pytest
import torch

import ffdas_amd
from ffdas_amd import _core

SCALAR_DTYPES = {"half": ffdas_amd.half, "float32": ffdas_amd.float32,
                 "float64": ffdas_amd.float64, "int16": ffdas_amd.int16, "int32": ffdas_amd.int32}
PACKED_DTYPES = {"complex64": ffdas_amd.complex64, "complex128": ffdas_amd.complex128,
                 "half2": ffdas_amd.half2, "float2": ffdas_amd.float2,
                 "double2": ffdas_amd.double2, "short2": ffdas_amd.short2, "int2": ffdas_amd.int2}
ALL_DTYPES = {**SCALAR_DTYPES, **PACKED_DTYPES}


@pytest.mark.parametrize("name, dt", ALL_DTYPES.items())
def test_dtype_name(name, dt):
    assert dt.name == name


@pytest.mark.parametrize("name, dt", ALL_DTYPES.items())
def test_dtype_repr_contains_name(name, dt):
    assert name in repr(dt)


def test_dtype_hashable():
    d = {ffdas_amd.float32: "f32", ffdas_amd.float64: "f64"}
    assert d[ffdas_amd.float32] == "f32"


def test_algorithm_values():
    assert ffdas_amd.Algorithm.DEFAULT.value == 0
    assert ffdas_amd.Algorithm.ALG1.value == 1


def test_interp_mode_values():
    assert ffdas_amd.InterpMode.NEAREST.value == 0
    assert ffdas_amd.InterpMode.LINEAR.value == 1


def test_view_float2_shape_cpu():
    v = ffdas_amd.view(torch.zeros((4, 3, 2)), ffdas_amd.float2)
    assert v.dtype == ffdas_amd.float2
    assert v.shape == (4, 3)


def test_view_preserves_base_cpu():
    base = torch.ones((8, 2))
    assert ffdas_amd.view(base, ffdas_amd.float2).base is base


def test_view_errors():
    with pytest.raises(ValueError, match="compatible"):
        ffdas_amd.view(torch.zeros((4, 2), dtype=torch.float64), ffdas_amd.float2)
    with pytest.raises(ValueError, match="trailing dimension of 2"):
        ffdas_amd.view(torch.zeros((4, 3)), ffdas_amd.float2)


_get_ptr = ctypes.PYFUNCTYPE(ctypes.c_void_p, ctypes.py_object, ctypes.c_char_p)(
    ("PyCapsule_GetPointer", ctypes.pythonapi))
_set_name = ctypes.PYFUNCTYPE(ctypes.c_int, ctypes.py_object, ctypes.c_char_p)(
    ("PyCapsule_SetName", ctypes.pythonapi))


def _capsule_tensor(cap):
    return ctypes.cast(_get_ptr(cap, b"dltensor"), ctypes.POINTER(_core._DLManagedTensor)).contents


@pytest.mark.parametrize("dt, base_dtype, code, bits, lanes", [
    (ffdas_amd.half2, torch.float16, 2, 16, 2), (ffdas_amd.float2, torch.float32, 2, 32, 2),
    (ffdas_amd.double2, torch.float64, 2, 64, 2), (ffdas_amd.short2, torch.int16, 0, 16, 2),
    (ffdas_amd.int2, torch.int32, 0, 32, 2), (ffdas_amd.bfloat162, torch.bfloat16, 4, 16, 2)])
def test_view_dlpack_capsule_fields(dt, base_dtype, code, bits, lanes):
    base = torch.zeros((5, 7, 2), dtype=base_dtype)[:, 1:6]     # strided view, (5, 5, 2)
    v = ffdas_amd.view(base, dt)
    cap = v.__dlpack__()
    mt = _capsule_tensor(cap)
    t = mt.dl_tensor
    assert (t.device.device_type, t.device.device_id) == v.__dlpack_device__() == (1, 0)
    assert t.data == base.data_ptr()
    assert (t.dtype.code, t.dtype.bits, t.dtype.lanes) == (code, bits, lanes)
    assert t.ndim == 2
    assert [t.shape[i] for i in range(2)] == [5, 5]
    assert [t.strides[i] for i in range(2)] == [7, 1]
    # a consumer renames the capsule and calls the deleter, which releases the export
    n = len(_core._dl_live)
    _set_name(cap, b"used_dltensor")
    mt.deleter(ctypes.pointer(mt))
    assert len(_core._dl_live) == n - 1
    del cap


def test_view_dlpack_unconsumed_capsule_released():
    import gc
    n = len(_core._dl_live)
    cap = ffdas_amd.view(torch.zeros((3, 2)), ffdas_amd.float2).__dlpack__()
    assert len(_core._dl_live) == n + 1
    del cap
    gc.collect()
    assert len(_core._dl_live) == n


def test_view_dlpack_torch_import_complex64():
    base = torch.randn(6, dtype=torch.complex64)
    v = ffdas_amd.view(base, ffdas_amd.complex64)
    t = torch.from_dlpack(v)
    assert t.dtype == torch.complex64 and t.data_ptr() == base.data_ptr()
    assert torch.equal(t, base)
