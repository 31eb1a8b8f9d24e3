"""Shared helpers for the tests (host side)."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)


def to_dev(w, dev, x_dtype=None):
    """Workload -> dict of torch tensors on dev (x optionally cast)."""
    import torch
    x = torch.from_numpy(np.ascontiguousarray(w.x)).to(dev)
    if x_dtype is not None:
        x = x.to(x_dtype)
    d = dict(x=x,
             srcpos=torch.from_numpy(w.srcpos).to(dev),
             dstpos=torch.from_numpy(w.dstpos).to(dev),
             offsets=torch.from_numpy(w.offsets).to(dev),
             weights=torch.from_numpy(w.weights).to(dev))
    if w.srcdir is not None:
        d["srcdir"] = torch.from_numpy(w.srcdir).to(dev)
    if w.sparse_indices is not None:
        d["sparse_indices"] = torch.from_numpy(w.sparse_indices).to(dev)
    if w.wavenum:
        d["wavenum"] = float(w.wavenum)
    return d


def oracle_args(w):
    kw = dict(srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices)
    return (w.x, w.srcpos, w.dstpos, w.offsets, w.weights), kw
