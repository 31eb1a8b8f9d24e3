"""ctypes binding of libffdas.so (the C-ABI in include/ffdas.h).

This is the Python side of the drop-in boundary: the reference binds the same C entry
points through nanobind (bindings/python/ffdas/_core/src/_ffdas.cpp:496-593) after loading
libffdas.so with ctypes (bindings/python/ffdas/__init__.py:9-90).  Loading fails loudly when
the library has not been built: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import enum
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_NAME = "libffdas.so"


class FfdasError(RuntimeError):
    """A non-success ffdas_error_t returned by the C library."""

    def __init__(self, code: int, where: str = ""):
        self.code = int(code)
        msg = error_string(code)
        super().__init__(f"{where}: {msg} (ffdas error {code})" if where else msg)


class DType(enum.IntEnum):
    """ffdas_datatype_t (include/ffdas.h)."""
    R_16F = 2
    C_16F = 6
    R_32F = 0
    C_32F = 4
    R_64F = 16
    C_64F = 17
    R_32I = 10
    C_32I = 11
    R_16I = 14
    C_16I = 15
    R_16BF = 18
    C_16BF = 19


class ComputeType(enum.IntEnum):
    """ffdas_compute_type_t; the reference's ComputeType (_ffdas.cpp:318-322) plus FP16BF."""
    DEFAULT = 0
    FP16 = 1
    FP32 = 2
    FP64 = 3
    BF16 = 4


class Algorithm(enum.IntEnum):
    """ffdas_alg_t; the reference's Algorithm (_ffdas.cpp:308-313)."""
    DEFAULT = 0
    ALG1 = 1
    ALG2 = 2
    ALG3 = 3
    ALG4 = 4


class InterpMode(enum.IntEnum):
    """ffdas_interp_mode_t; the reference's InterpMode (_ffdas.cpp:315-317)."""
    NEAREST = 0
    LINEAR = 1


class DasProblem(ctypes.Structure):
    """ffdas_das_problem_t (include/ffdas.h)."""
    _fields_ = [("srcpos", ctypes.c_void_p), ("srcdir", ctypes.c_void_p),
                ("wavenum", ctypes.c_float), ("scale", ctypes.c_float),
                ("x_desc", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("dstpos", ctypes.c_void_p), ("offsets", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("sparse_count", ctypes.c_int),
                ("sparse_indices", ctypes.c_void_p), ("beta", ctypes.c_void_p),
                ("out_desc", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class DasTuning(ctypes.Structure):
    """ffdas_das_tuning_t (include/ffdas.h)."""
    _fields_ = [("groups", ctypes.c_int), ("fi_frames", ctypes.c_int),
                ("tile_h", ctypes.c_int), ("tile_w", ctypes.c_int),
                ("param_groups", ctypes.c_int), ("bwdx_frames", ctypes.c_int),
                ("bwdx_depth_lanes", ctypes.c_int), ("bwdx_threads", ctypes.c_int),
                ("forward_ms", ctypes.c_float), ("backward_ms", ctypes.c_float),
                ("default_forward_ms", ctypes.c_float), ("default_backward_ms", ctypes.c_float),
                ("candidates", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def _lib_path() -> Path:
    env = os.environ.get("FFDAS_LIB_DIR")
    return Path(env) / LIB_NAME if env else _PKG / LIB_NAME


_lib = None


def lib():
    """The loaded libffdas.so (raises ImportError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = _lib_path()
    if not path.exists():
        raise ImportError(
            f"{path} not found: build the HIP library first "
            f"(python -m ffdas_amd.build, or __graft_entry__.build())")
    L = ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)
    vp, i32, i64, f32, u = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float,
                            ctypes.c_uintptr_t if hasattr(ctypes, "c_uintptr_t") else ctypes.c_size_t)
    P = ctypes.POINTER
    sig = {
        "ffdas_type_size": (ctypes.c_size_t, [i32]),
        "ffdas_error_string": (ctypes.c_char_p, [i32]),
        "ffdas_version": (ctypes.c_char_p, []),
        "ffdas_create_tensor_desc": (i32, [P(vp), i64, P(i64), P(i64), i32]),
        "ffdas_destroy_tensor_desc": (i32, [vp]),
        "ffdas_create": (i32, [P(vp)]),
        "ffdas_destroy": (i32, [vp]),
        "ffdas_set_stream": (i32, [vp, u]),
        "ffdas_get_stream": (i32, [vp, P(u)]),
        "ffdas_contiguous_copy": (i32, [vp, vp, vp, vp, vp]),
        "ffdas_das": (i32, [vp, vp, vp, f32, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32]),
        "ffdas_das_noalloc": (i32, [vp, vp, vp, f32, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32]),
        "ffdas_das_sparse": (i32, [vp, vp, vp, f32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32,
                                   i32]),
        "ffdas_das_ex": (i32, [vp, vp, vp, f32, f32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32,
                               i32]),
        "ffdas_das_backward": (i32, [vp, vp, vp, f32, f32, vp, vp, vp, vp, vp, i32, vp, vp, vp,
                                     vp, vp, vp, vp, vp, i32, i32]),
        "ffdas_das_validate": (i32, [vp, vp, vp, vp, vp, vp, i32, vp, i32, i32]),
        "ffdas_device_count": (i32, [P(i32)]),
        "ffdas_device_get": (i32, [P(i32)]),
        "ffdas_device_set": (i32, [i32]),
        "ffdas_event_create": (i32, [P(u)]),
        "ffdas_event_destroy": (i32, [u]),
        "ffdas_event_record": (i32, [vp, u]),
        "ffdas_event_synchronize": (i32, [u]),
        "ffdas_event_elapsed_time": (i32, [u, u, P(f32)]),
        "ffdas_das_autotune": (i32, [vp, vp, vp, f32, f32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp,
                                     vp, i32, i32, i32, P(DasTuning)]),
        "ffdas_das_batched": (i32, [vp, i32, P(DasProblem), i32, i32]),
        "ffdas_set_reuse_x": (i32, [vp, i32]),
        "ffdas_tune_save": (i32, [ctypes.c_char_p]),
        "ffdas_tune_load": (i32, [ctypes.c_char_p]),
        "ffdas_tune_clear": (i32, []),
        "ffdas_tune_count": (i32, []),
        "ffdas_truncate_rank": (i32, [vp, vp, vp, i64, i64, vp, vp]),
        "ffdas_create_interpolation_plan": (i32, [vp, P(vp), i64, i64, i64, vp, i32]),
        "ffdas_interpolation_preprocess": (i32, [vp, vp, i64, vp]),
        "ffdas_interpolation": (i32, [vp, vp, i64, vp, vp, vp, vp, vp]),
        "ffdas_destroy_interpolation_plan": (i32, [vp, vp]),
        "ffdas_greens": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "ffdas_profile_enable": (i32, [i32]),
        "ffdas_profile_reset": (i32, []),
        "ffdas_profile_query": (i32, [ctypes.c_char_p, P(ctypes.c_double), P(i64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def error_string(code: int) -> str:
    return lib().ffdas_error_string(int(code)).decode()


def check(code: int, where: str = "") -> None:
    if code != 0:
        raise FfdasError(code, where)


def exported_symbols() -> list[str]:
    """Names declared with FFDAS_API in include/ffdas.h."""
    import re
    hdr = (_PKG.parent / "include" / "ffdas.h").read_text()
    return re.findall(r"FFDAS_API\s+[^;]*?\b(ffdas_\w+)\s*\(", hdr)


class TensorDesc:
    """RAII wrapper of an ffdas_tensor_desc_t (dims/strides in elements)."""

    def __init__(self, dims, strides, dtype: int):
        L = lib()
        n = len(dims)
        d = (ctypes.c_int64 * n)(*[int(v) for v in dims])
        s = (ctypes.c_int64 * n)(*[int(v) for v in strides]) if strides is not None else None
        self.ptr = ctypes.c_void_p()
        check(L.ffdas_create_tensor_desc(ctypes.byref(self.ptr), n, d, s, int(dtype)),
              "ffdas_create_tensor_desc")

    def __del__(self):
        try:
            if self.ptr:
                lib().ffdas_destroy_tensor_desc(self.ptr)
                self.ptr = ctypes.c_void_p()
        except Exception:
            pass
