"""LDS atomic throughput of the device (the on-chip ceiling of the grad_x scatter).

Loads ffdas_amd/calib/libffdas_calib.so (built with the package by ffdas_amd/build.py) and
times its probe kernel: every lane adds into LDS at a fixed per-lane word offset, 16 atomics
per loop step, so the measured rate is that of one address pattern.  With distinct banks
this is the conflict-free peak the grad_x adjoint is set against.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

_LIB = Path(__file__).resolve().parents[1] / "calib" / "libffdas_calib.so"
_lib = None

OPS = {"u64": 0, "u32": 1, "f32": 2}


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise ImportError(f"{_LIB} not found: build the package first (ffdas_amd.build)")
        L = ctypes.CDLL(str(_LIB))
        L.ffdas_calib_lds_atomics.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_float)]
        L.ffdas_calib_lds_atomics.restype = ctypes.c_int
        L.ffdas_calib_device_info.argtypes = [ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]
        L.ffdas_calib_device_info.restype = ctypes.c_int
        _lib = L
    return _lib


def device_info() -> tuple[int, float]:
    """(compute units, peak engine clock in GHz) of the current device."""
    cus, khz = ctypes.c_int(), ctypes.c_int()
    rc = _load().ffdas_calib_device_info(ctypes.byref(cus), ctypes.byref(khz))
    if rc != 0:
        raise RuntimeError(f"ffdas_calib_device_info: hip error {rc}")
    return cus.value, khz.value * 1e-6


def lds_atomic_rate(op: str = "u64", pattern=None, reps: int = 256, blocks: int | None = None,
                    threads: int = 512) -> dict:
    """Lane-ops per clock per CU of LDS atomics `op` ("u64", "u32" or "f32") when lane l of
    every wave adds at word pattern[l] (default: lane l -> word l, distinct banks; -1: the
    lane stays idle and is not counted)."""
    cus, ghz = device_info()
    pat = list(range(64)) if pattern is None else [int(v) for v in pattern]
    if len(pat) != 64:
        raise ValueError("pattern must give one word offset per lane (64)")
    blocks = blocks or 8 * cus
    ms = ctypes.c_float()
    arr = (ctypes.c_int * 64)(*pat)
    rc = _load().ffdas_calib_lds_atomics(OPS[op], arr, reps, blocks, threads, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"ffdas_calib_lds_atomics: hip error {rc}")
    active = sum(1 for v in pat if v >= 0)
    lane_ops = blocks * (threads // 64) * active * reps * 16
    rate = lane_ops / (ms.value * 1e-3) / (ghz * 1e9 * cus)
    return {"op": op, "ms": ms.value, "lane_ops_per_clk_cu": rate, "cus": cus,
            "clock_ghz": ghz, "blocks": blocks, "threads": threads}
