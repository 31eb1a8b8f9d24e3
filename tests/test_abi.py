"""C-ABI checks that need no GPU: the library loads, exports every symbol include/ffdas.h
declares, keeps the reference's enum values, and returns the reference's error codes for
every argument-validation branch (via the host-only ffdas_das_validate)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from ffdas_amd import _lib
from ffdas_amd._lib import Algorithm, ComputeType, DType, TensorDesc

ROOT = Path(__file__).resolve().parents[1]
HDR = (ROOT / "include" / "ffdas.h").read_text()

E = dict(SUCCESS=0, INVALID_ARGUMENT=1, INVALID_DIMS=2, UNSUPPORTED_TYPE=3, NULL_POINTER=4,
         INVALID_HANDLE=5, NON_CONTIGUOUS=6, DIMS_TOO_LARGE=7, DEVICE_NOT_AVAILABLE=22,
         DEVICE_MISMATCH=23)


def L():
    return _lib.lib()


def test_library_exports_every_declared_symbol():
    declared = set(_lib.exported_symbols())
    assert {"ffdas_das", "ffdas_das_sparse", "ffdas_create", "ffdas_das_backward"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "ffdas_amd" / "libffdas.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared <= exported, declared - exported
    for name in declared:
        assert hasattr(L(), name)


def test_enum_values_match_reference():
    """include/ffdas.h:19-90 of the reference."""
    ref = {"FFDAS_R_16F": 2, "FFDAS_C_16F": 6, "FFDAS_R_32F": 0, "FFDAS_C_32F": 4,
           "FFDAS_R_64F": 16, "FFDAS_C_64F": 17, "FFDAS_R_32I": 10, "FFDAS_C_32I": 11,
           "FFDAS_R_16I": 14, "FFDAS_C_16I": 15, "FFDAS_COMPUTE_DEFAULT": 0,
           "FFDAS_COMPUTE_16F": 1, "FFDAS_COMPUTE_32F": 2, "FFDAS_COMPUTE_64F": 3,
           "FFDAS_ALG_DEFAULT": 0, "FFDAS_ALG1": 1, "FFDAS_ALG2": 2, "FFDAS_ALG3": 3,
           "FFDAS_ALG4": 4, "FFDAS_SUCCESS": 0, "FFDAS_ERROR_INVALID_ARGUMENT": 1,
           "FFDAS_ERROR_INVALID_DIMS": 2, "FFDAS_ERROR_UNSUPPORTED_TYPE": 3,
           "FFDAS_ERROR_NULL_POINTER": 4, "FFDAS_ERROR_INVALID_HANDLE": 5,
           "FFDAS_ERROR_DEVICE_NOT_AVAILABLE": 22, "FFDAS_ERROR_DEVICE_MISMATCH": 23,
           "FFDAS_ERROR_CUDA_RUNTIME": 31, "FFDAS_ERROR_UNKNOWN": 101}
    for name, val in ref.items():
        m = re.search(rf"\b{name}\s*=\s*(\d+)", HDR)
        assert m and int(m.group(1)) == val, name
    assert DType.R_16F == 2 and DType.C_64F == 17 and ComputeType.FP16 == 1
    assert Algorithm.ALG4 == 4


def test_type_sizes_and_error_strings():
    f = L().ffdas_type_size
    assert [f(int(t)) for t in (DType.R_16F, DType.C_16F, DType.R_32F, DType.C_32F, DType.R_64F,
                                DType.C_64F, DType.R_16BF, DType.C_16BF)] == [2, 4, 4, 8, 8, 16,
                                                                               2, 4]
    assert f(99) == 0
    assert _lib.error_string(0) == "Success"
    assert _lib.error_string(3) == "Unsupported data type"
    assert _lib.error_string(12345) == "Unrecognized error code"


def test_tensor_desc_errors():
    p = ctypes.c_void_p()
    dims = (ctypes.c_int64 * 2)(3, 4)
    assert L().ffdas_create_tensor_desc(ctypes.byref(p), 0, dims, None, 0) == E["INVALID_ARGUMENT"]
    assert L().ffdas_create_tensor_desc(None, 2, dims, None, 0) == E["INVALID_ARGUMENT"]
    assert L().ffdas_create_tensor_desc(ctypes.byref(p), 2, dims, None, 99) == E["UNSUPPORTED_TYPE"]
    assert L().ffdas_create_tensor_desc(ctypes.byref(p), 2, dims, None, 0) == 0
    assert L().ffdas_destroy_tensor_desc(p) == 0
    assert L().ffdas_destroy_tensor_desc(None) == 0


def test_handle_functions_reject_null():
    assert L().ffdas_destroy(None) == E["INVALID_HANDLE"]
    assert L().ffdas_set_stream(None, 0) == E["INVALID_HANDLE"]
    assert L().ffdas_das(None, None, None, 0.0, None, None, None, None, None, None, None, None,
                         0, 0) == E["INVALID_HANDLE"]
    assert L().ffdas_event_record(None, 1) == E["INVALID_HANDLE"]
    assert L().ffdas_event_synchronize(0) == E["INVALID_ARGUMENT"]


def test_create_without_gpu_reports_not_available():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert L().ffdas_create(ctypes.byref(h)) == E["DEVICE_NOT_AVAILABLE"]
    c = ctypes.c_int(-1)
    assert L().ffdas_device_count(ctypes.byref(c)) == 0 and c.value == 0


# ---- validation branches (host only) ----------------------------------------------------

NZ = ctypes.c_void_p(16)   # non-null dummy device pointer (never dereferenced by validate)


def validate(x_dims, out_dims, x_dtype=DType.R_32F, out_dtype=DType.R_32F, x_strides=None,
             out_strides=None, sparse_count=0, sparse=False, ct=ComputeType.DEFAULT,
             alg=Algorithm.DEFAULT, null_geom=False):
    xd = TensorDesc(x_dims, x_strides, x_dtype)
    od = TensorDesc(out_dims, out_strides, out_dtype)
    g = None if null_geom else NZ
    return L().ffdas_das_validate(xd.ptr, od.ptr, g, NZ, NZ, NZ, sparse_count,
                                  NZ if sparse else None, int(ct), int(alg))


def test_validate_accepts_supported_pairs():
    for xt, ot in [(DType.R_16F, DType.R_32F), (DType.R_32F, DType.R_32F),
                   (DType.R_64F, DType.R_64F), (DType.C_16F, DType.C_32F),
                   (DType.C_32F, DType.C_32F), (DType.C_64F, DType.C_64F),
                   (DType.R_16BF, DType.R_32F), (DType.C_16BF, DType.C_32F)]:
        assert validate((2, 4, 3, 64), (2, 10), xt, ot) == 0, (xt, ot)


def test_validate_reference_error_branches():
    # unsupported dtype pair (src/das/das.cu:82)
    assert validate((4, 3, 64), (10,), DType.R_32F, DType.R_64F) == E["UNSUPPORTED_TYPE"]
    assert validate((4, 3, 64), (10,), DType.R_16F, DType.R_16F) == E["UNSUPPORTED_TYPE"]
    assert validate((4, 3, 64), (10,), DType.R_32I, DType.R_32F) == E["UNSUPPORTED_TYPE"]
    # FP64 with ALG2/ALG4 (das_impl.cuh:37-40)
    assert validate((4, 3, 64), (10,), DType.R_64F, DType.R_64F,
                    alg=Algorithm.ALG2) == E["UNSUPPORTED_TYPE"]
    # ALG3 (das_impl.cuh:55)
    assert validate((4, 3, 64), (10,), alg=Algorithm.ALG3) == E["INVALID_ARGUMENT"]
    assert validate((4, 3, 64), (10,), alg=Algorithm.ALG2) == 0
    # x must be 3-D or 4-D (das_utils.h:39-40)
    assert validate((3, 64), (10,)) == E["INVALID_DIMS"]
    assert validate((1, 2, 4, 3, 64), (1, 10)) == E["INVALID_DIMS"]
    # batch mismatch (das_utils.h:52-53)
    assert validate((2, 4, 3, 64), (3, 10)) == E["INVALID_DIMS"]
    assert validate((2, 4, 3, 64), (10,)) == E["INVALID_DIMS"]
    # sample stride != 1 (das_utils.h:66-67)
    assert validate((4, 3, 64), (10,), x_strides=(384, 128, 2)) == E["INVALID_ARGUMENT"]
    # sparse: count <= 0 (das_alg1.cu:267), only DEFAULT/ALG1 (das_sparse_impl.cuh:38-45)
    assert validate((4, 3, 64), (10,), sparse=True, sparse_count=0) == E["INVALID_ARGUMENT"]
    assert validate((4, 3, 64), (10,), sparse=True, sparse_count=2) == 0
    assert validate((4, 3, 64), (10,), sparse=True, sparse_count=2,
                    alg=Algorithm.ALG2) == E["INVALID_ARGUMENT"]
    # COMPUTE_16F only converts float data (das_alg1.cu:172-178)
    assert validate((4, 3, 64), (10,), ct=ComputeType.FP16) == 0
    assert validate((4, 3, 64), (10,), DType.R_16F, DType.R_32F,
                    ct=ComputeType.FP16) == E["UNSUPPORTED_TYPE"]
    assert validate((4, 3, 64), (10,), DType.R_64F, DType.R_64F,
                    ct=ComputeType.FP16) == E["UNSUPPORTED_TYPE"]
    # COMPUTE_32F on double data: float storage, double output (das_alg1.cu:179-184)
    assert validate((4, 3, 64), (10,), DType.R_64F, DType.R_64F, ct=ComputeType.FP32) == 0


def test_validate_extensions():
    # null geometry pointers are rejected instead of faulting
    assert validate((4, 3, 64), (10,), null_geom=True) == E["NULL_POINTER"]
    # non-contiguous output
    assert validate((2, 4, 3, 64), (2, 5, 4), out_strides=(20, 1, 5)) == E["NON_CONTIGUOUS"]
    # strided batch of the output is accepted
    assert validate((2, 4, 3, 64), (2, 10), out_strides=(16, 1)) == 0
    # negative strides rejected
    assert validate((2, 4, 3, 64), (2, 10), x_strides=(-768, 192, 64, 1)) == E["INVALID_ARGUMENT"]
    # one frame beyond 2^31 elements
    assert validate((1, 70000, 64, 1024), (1, 10), DType.R_16F) == E["INVALID_DIMS"]
    # a whole tensor beyond 2^31 elements is fine when one frame is not (64-bit frame offsets)
    assert validate((300, 1024, 64, 128), (300, 10), DType.R_16F) == 0
    assert L().ffdas_das_validate(None, None, NZ, NZ, NZ, NZ, 0, None, 0, 0) == E["NULL_POINTER"]


def test_tuning_cache_roundtrip(tmp_path):
    """ffdas_tune_load / _save / _clear / _count are host-only: a file written by hand is
    loaded, saved back unchanged, and rejected when malformed."""
    import ffdas_amd
    ffdas_amd.clear_tuning()
    assert ffdas_amd.tuning_entries() == 0
    src = tmp_path / "tune.txt"
    lines = ["gfx950|x=f32|k=f16|c=0|a64=0|sp=0|C=128|T=11|S=2048|B=16|N=256x256|dir=0 "
             "8 16 8 8 4 16 4 1024",
             "gfx950|x=f32|k=f32|c=1|a64=0|sp=8|C=64|T=64|S=2048|B=200|N=128x128|dir=1 "
             "4 8 16 4 2 8 3 512"]
    src.write_text("# comment\n" + "\n".join(lines) + "\n")
    ffdas_amd.load_tuning(src)
    assert ffdas_amd.tuning_entries() == 2
    out = tmp_path / "out.txt"
    ffdas_amd.save_tuning(out)
    assert sorted(out.read_text().split("\n")[:-1]) == sorted(lines)
    bad = tmp_path / "bad.txt"
    bad.write_text("key 1 2 3\n")
    with pytest.raises(ffdas_amd.FfdasError):
        ffdas_amd.load_tuning(bad)
    with pytest.raises(ffdas_amd.FfdasError):
        ffdas_amd.load_tuning(tmp_path / "missing.txt")
    ffdas_amd.clear_tuning()
    assert ffdas_amd.tuning_entries() == 0


def test_descriptor_stride_overflow_is_refused():
    """Contiguous strides that do not fit in 64 bits are refused with
    FFDAS_ERROR_DIMS_TOO_LARGE instead of overflowing."""
    d = ctypes.c_void_p()
    dims = (ctypes.c_int64 * 4)(1 << 40, 1 << 40, 1 << 40, 9)     # stride of dim 0: 9 * 2^80
    assert L().ffdas_create_tensor_desc(ctypes.byref(d), 4, dims, None, int(DType.R_32F)) == 7
    dims = (ctypes.c_int64 * 3)(1 << 20, 1 << 20, 9)
    assert L().ffdas_create_tensor_desc(ctypes.byref(d), 3, dims, None, int(DType.R_32F)) == 0
    L().ffdas_destroy_tensor_desc(d)
