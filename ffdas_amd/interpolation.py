"""Structured-grid interpolation: ``Interpolator`` and ``interpolate`` with the reference's
Python signatures and errors (bindings/python/ffdas/interpolation.py:15-190), on the
ffdas_*interpolation* entry points of include/ffdas.h."""
from __future__ import annotations

import ctypes

import torch

from ._core import as_tensor, check_device_tensor, element_code, handle_for
from ._lib import InterpMode, TensorDesc, check, lib

__all__ = ["Interpolator", "interpolate"]

_MODES = {"nearest": 0, "linear": 1}
_SUPPORTED = (torch.float32, torch.complex64, torch.float64, torch.complex128)


class Interpolator:
    """Structured 3D grid interpolation with a reusable plan.

    Args:
        gridpos: grid vertex positions, shape (nz, ny, nx, 3) (converted to float32).
        mode: "nearest" or "linear" (or :class:`InterpMode`).
    """

    def __init__(self, gridpos, mode="linear"):
        self._plan = None
        gridpos = as_tensor(gridpos)
        if gridpos.ndim != 4 or gridpos.shape[3] != 3:
            raise ValueError("gridpos must have shape (nz, ny, nx, 3)")
        if isinstance(mode, InterpMode):
            mode = mode.name.lower()
        if mode not in _MODES:
            raise ValueError(f"invalid interpolation mode '{mode}' "
                             f"(must be one of {tuple(_MODES.keys())})")
        check_device_tensor(gridpos, "gridpos")
        self.shape = tuple(int(d) for d in gridpos.shape[:3])
        self.mode = mode
        self.device = gridpos.device
        g = gridpos.to(torch.float32).contiguous()
        nz, ny, nx = self.shape
        plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            self._h = handle_for(g)
            check(lib().ffdas_create_interpolation_plan(self._h.h, ctypes.byref(plan), nx, ny, nz,
                                                        g.data_ptr(), _MODES[mode]),
                  "ffdas_create_interpolation_plan")
        self._plan = plan

    def __call__(self, x, querypos, *, fill=None, out=None, preprocess: bool = False):
        """Values at the query points: shape querypos.shape[:-1], with x's leading batch axis
        when x is 4-D; points outside the grid get fill (default 0)."""
        if self._plan is None:
            raise RuntimeError("Interpolator has been destroyed")
        x = as_tensor(x)
        if x.ndim < 3 or x.ndim > 4:
            raise ValueError(f"x must be three or four-dimensional (got {tuple(x.shape)})")
        if tuple(x.shape[-3:]) != self.shape:
            raise ValueError("x must have the same leading dimensions as the configured "
                             f"interpolation points {tuple(x.shape[-3:])} != {self.shape})")
        if x.dtype not in _SUPPORTED:
            raise TypeError(f"unsupported dtype {x.dtype} (float32, complex64, float64 or "
                            "complex128)")
        check_device_tensor(x, "x", self.device)
        if preprocess:
            self.preprocess(querypos)
        q = as_tensor(querypos)
        check_device_tensor(q, "querypos", self.device)
        if q.shape[-1] != 3:
            raise ValueError("querypos must have shape (..., 3)")
        q = q.to(torch.float32).contiguous()
        nq = q.numel() // 3
        if fill is None:
            fill = 0.0
        fill_t = torch.tensor([fill], dtype=x.dtype)          # host copy of one element
        out_shape = tuple(q.shape[:-1])
        if x.ndim == 4:
            out_shape = (int(x.shape[0]),) + out_shape
        if out is None:
            out = torch.empty(out_shape, dtype=x.dtype, device=x.device)
        elif tuple(out.shape) != out_shape:
            raise ValueError(f"incorrect dimensions for out: expected {out_shape}, "
                             f"got {tuple(out.shape)})")
        else:
            check_device_tensor(out, "out", self.device)
            if out.dtype != x.dtype or not out.is_contiguous():
                raise ValueError("out must be a contiguous tensor of x's dtype")
        xc = x.contiguous()
        xd = TensorDesc(tuple(xc.shape), tuple(xc.stride()), element_code(xc))
        with torch.cuda.device(self.device):
            h = handle_for(xc)
            check(lib().ffdas_interpolation(h.h, self._plan, nq, q.data_ptr(), xd.ptr,
                                            xc.data_ptr(), out.data_ptr(), fill_t.data_ptr()),
                  "ffdas_interpolation")
        return out

    def preprocess(self, querypos) -> None:
        """Locate the given query points once; later calls with as many points reuse it."""
        if self._plan is None:
            raise RuntimeError("Interpolator has been destroyed")
        q = as_tensor(querypos)
        check_device_tensor(q, "querypos", self.device)
        q = q.to(torch.float32).contiguous()
        with torch.cuda.device(self.device):
            h = handle_for(q)
            check(lib().ffdas_interpolation_preprocess(h.h, self._plan, q.numel() // 3,
                                                       q.data_ptr()),
                  "ffdas_interpolation_preprocess")

    def close(self) -> None:
        """Release the plan and its device memory."""
        if self._plan is not None:
            lib().ffdas_destroy_interpolation_plan(self._h.h, self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def interpolate(gridpos, x, querypos, *, mode: str = "linear", fill=None, out=None):
    """One-shot ``Interpolator(gridpos, mode)(x, querypos, fill=fill, out=out)``."""
    interp = Interpolator(gridpos, mode)
    try:
        return interp(x, querypos, fill=fill, out=out)
    finally:
        interp.close()
