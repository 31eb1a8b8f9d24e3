"""Transmit-geometry helpers: the step before delay-and-sum that turns positions into
transmit delays (offsets) and apodization weights.  Same functions, arguments and results
as the reference's ffdas.utils.spatial (bindings/python/ffdas/utils/spatial.py:10-50);
they accept torch tensors (any device) or numpy arrays and return the same kind."""
from __future__ import annotations

import numpy as np

__all__ = ["cdist", "rect_dist", "angle"]


def _is_torch(*xs) -> bool:
    try:
        import torch
    except ImportError:          # pragma: no cover - torch is part of this stack
        return False
    return any(isinstance(x, torch.Tensor) for x in xs)


def cdist(a, b):
    """Euclidean distance from each point of a to each point of b:
    a (...A, D), b (...B, D) -> (*a.shape[:-1], *b.shape[:-1])."""
    na, nb, d = a.ndim - 1, b.ndim - 1, a.shape[-1]
    a = a.reshape(tuple(a.shape[:-1]) + (1,) * nb + (d,))
    b = b.reshape((1,) * na + tuple(b.shape))
    diff = a - b
    return (diff * diff).sum(-1) ** 0.5


def rect_dist(points, size):
    """Distance from 3-D points (..., 3) to the axis-aligned rectangle of full width and
    height size = (w, h) in the z = 0 plane, centred at the origin."""
    if _is_torch(points, size):
        import torch
        size = torch.as_tensor(size, dtype=points.dtype, device=points.device)
        clip = lambda v: torch.clamp(v, min=0)
    else:
        size = np.asarray(size, dtype=np.asarray(points).dtype)
        clip = lambda v: np.clip(v, 0, None)
    half = size / 2
    dx = clip(abs(points[..., 0]) - half[0])
    dy = clip(abs(points[..., 1]) - half[1])
    dz = abs(points[..., 2])
    return (dx * dx + dy * dy + dz * dz) ** 0.5


def angle(a, b, eps=1e-7):
    """Angle in radians between the vectors a and b (..., D) -> (...)."""
    dot = (a * b).sum(-1)
    na = ((a * a).sum(-1)) ** 0.5
    nb = ((b * b).sum(-1)) ** 0.5
    c = dot / (na * nb + eps)
    if _is_torch(a, b):
        import torch
        return torch.acos(torch.clamp(c, -1.0, 1.0))
    return np.arccos(np.clip(c, -1.0, 1.0))
