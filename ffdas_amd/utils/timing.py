"""GPU timer on the library's stream (reference bindings/python/ffdas/utils/timing.py:5-62).

Events are hipEvents recorded with ffdas_event_record on the handle of the current device,
which is bound to torch's current stream.
"""
from __future__ import annotations

import ctypes

import torch

from .._core import get_library_handle
from .._lib import check, lib


class Timer:
    """Context manager measuring GPU time between start() and stop().

    Example::

        with Timer() as t:
            out = ffdas_amd.das(rf, srcpos, dstpos, offsets, weights)
        print(f"{t.elapsed_ms():.3f} ms")
    """

    def __init__(self):
        self._handle = get_library_handle()
        self._start = ctypes.c_size_t()
        self._stop = ctypes.c_size_t()
        check(lib().ffdas_event_create(ctypes.byref(self._start)), "ffdas_event_create")
        check(lib().ffdas_event_create(ctypes.byref(self._stop)), "ffdas_event_create")
        self._running = False
        self._elapsed = None

    def _record(self, ev):
        self._handle.set_stream(torch.cuda.current_stream().cuda_stream)
        check(lib().ffdas_event_record(self._handle.h, ev.value), "ffdas_event_record")

    def start(self):
        if self._running:
            raise RuntimeError("Timer is already running")
        self._running = True
        self._record(self._start)

    def stop(self):
        if not self._running:
            return
        self._record(self._stop)
        check(lib().ffdas_event_synchronize(self._stop.value), "ffdas_event_synchronize")
        ms = ctypes.c_float()
        check(lib().ffdas_event_elapsed_time(self._start.value, self._stop.value,
                                             ctypes.byref(ms)), "ffdas_event_elapsed_time")
        self._elapsed = float(ms.value)
        self._running = False

    def __enter__(self):
        self.start()
        return self

    def __exit__(self, *exc):
        self.stop()
        return False

    def elapsed_ms(self) -> float:
        if self._elapsed is None:
            raise RuntimeError("Timer has not been stopped yet")
        return self._elapsed

    def __del__(self):
        try:
            for ev in (self._start, self._stop):
                if ev.value:
                    lib().ffdas_event_destroy(ev.value)
                    ev.value = 0
        except Exception:
            pass
