"""Per-kernel GPU timing of the library (ffdas_profile_* in include/ffdas.h).

While enabled, every kernel the library launches is bracketed by hipEvents recorded on the
stream it is launched on; :func:`kernel_stats` returns the average duration per kernel.
"""
from __future__ import annotations

import ctypes
from contextlib import contextmanager

from .._lib import check, lib

KERNELS = ("das_fwd", "das_fwd_fi", "das_bwd_param", "das_bwd_param_fi", "das_bwd_x",
           "das_bwd_x_plain", "das_fwd_fi_batched", "relayout_fi", "cast_rows",
           "transpose_gout")


def enable(on: bool = True) -> None:
    check(lib().ffdas_profile_enable(1 if on else 0), "ffdas_profile_enable")


def reset() -> None:
    check(lib().ffdas_profile_reset(), "ffdas_profile_reset")


def query(name: str):
    """(total_ms, launches) for one kernel name."""
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().ffdas_profile_query(name.encode(), ctypes.byref(ms), ctypes.byref(n)),
          "ffdas_profile_query")
    return ms.value, n.value


def kernel_stats() -> dict:
    """{name: {"total_ms", "launches", "avg_ms"}} for every kernel launched since reset()."""
    out = {}
    for k in KERNELS:
        ms, n = query(k)
        if n:
            out[k] = {"total_ms": ms, "launches": n, "avg_ms": ms / n}
    return out


@contextmanager
def profiled():
    reset()
    enable(True)
    try:
        yield
    finally:
        enable(False)
