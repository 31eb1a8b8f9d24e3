"""Independent numpy FP64 restatement of DAS (TEST INFRASTRUCTURE ONLY).

Written separately from das_oracle.cpp, from the documented definition
(docs/api/das.md:5,37-48,90-100; docs/conventions.md:9-35) and the reference kernel's
arithmetic (src/das/das_alg1_kernels.cuh:58-99): receive radius
sqrt(max(|d|^2, 1e-7)), transmit delay added, floor/frac, zero outside 0 <= k < S-1,
strict directivity cone, complex weight w * exp(i * wavenum * tau), linear interpolation.
It broadcasts over (slot, channel, voxel) arrays instead of looping, so a shared
indexing mistake with the C++ oracle is unlikely.  Only for small inputs.
"""
from __future__ import annotations

import numpy as np

RADIUS_SQ_MIN = float(np.float32(1e-7))  # das_alg1_kernels.cuh:64, a float there


def taps(srcpos, dstpos, offsets, srcdir=None, scale=1.0, n_samples=None, sparse_indices=None,
         n_seq=None):
    """Return (valid, k, t, tau, r, seq) arrays of shape (K, C, N)."""
    srcpos = np.asarray(srcpos, np.float64)                      # (C,3)
    dst = np.asarray(dstpos, np.float64).reshape(-1, 3)          # (N,3)
    N = dst.shape[0]
    ofs = np.asarray(offsets, np.float64).reshape(-1, N)         # (K,N)
    K = ofs.shape[0]
    d = dst[None, :, :] - srcpos[:, None, :]                     # (C,N,3)
    r2 = np.maximum(np.einsum("cnk,cnk->cn", d, d), RADIUS_SQ_MIN)
    r = np.sqrt(r2)                                              # (C,N)
    tau = scale * (ofs[:, None, :] + r[None, :, :])              # (K,C,N)
    kf = np.floor(tau)
    t = tau - kf
    valid = (kf >= 0) & (kf < n_samples - 1)
    if sparse_indices is not None:
        seq = np.asarray(sparse_indices, np.int64).reshape(K, N)
        seq = np.broadcast_to(seq[:, None, :], tau.shape)
        valid &= (seq >= 0) & (seq < n_seq)
    else:
        seq = np.broadcast_to(np.arange(K)[:, None, None], tau.shape)
    if srcdir is not None:
        sd = np.asarray(srcdir, np.float64)                      # (C,4)
        proj = np.einsum("cnk,ck->cn", d, sd[:, :3]) / r
        valid &= (proj > sd[:, 3:4])[None, :, :]
    k = np.where(valid, kf, 0).astype(np.int64)
    return valid, k, t, tau, r, seq


def das(x, srcpos, dstpos, offsets, weights, *, srcdir=None, wavenum=0.0, sparse_indices=None,
        scale=1.0, channels_trailing=False):
    x = np.asarray(x)
    batched = x.ndim == 4
    if not batched:
        x = x[None]
    if channels_trailing:
        x = np.swapaxes(x, 1, 2)
    x = x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)
    B, C, T, S = x.shape
    spatial = np.asarray(dstpos).shape[:-1]
    N = int(np.prod(spatial))
    w = np.asarray(weights, np.float64).reshape(-1, N)
    valid, k, t, tau, r, seq = taps(srcpos, dstpos, offsets, srcdir, scale, S, sparse_indices, T)
    K = w.shape[0]
    ch = np.broadcast_to(np.arange(C)[None, :, None], k.shape)
    seqc = np.where(valid, seq, 0)
    x0 = x[:, ch, seqc, k]                                       # (B,K,C,N)
    x1 = x[:, ch, seqc, np.minimum(k + 1, S - 1)]
    interp = x0 + t[None] * (x1 - x0)
    coef = np.where(valid, w[:, None, :], 0.0)
    if np.iscomplexobj(x):
        coef = coef * np.exp(1j * float(np.float32(wavenum)) * tau)
    out = np.sum(coef[None] * interp, axis=(1, 2))               # (B,N)
    out = out.reshape((B,) + tuple(spatial))
    return out if batched else out[0]


def das_adjoint_x(x_shape, srcpos, dstpos, offsets, weights, grad_out, *, srcdir=None,
                  wavenum=0.0, sparse_indices=None, scale=1.0, is_complex=False):
    """Transpose of das with respect to x (numpy scatter with np.add.at)."""
    batched = len(x_shape) == 4
    B, C, T, S = x_shape if batched else (1,) + tuple(x_shape)
    spatial = np.asarray(dstpos).shape[:-1]
    N = int(np.prod(spatial))
    w = np.asarray(weights, np.float64).reshape(-1, N)
    valid, k, t, tau, r, seq = taps(srcpos, dstpos, offsets, srcdir, scale, S, sparse_indices, T)
    coef = np.where(valid, w[:, None, :], 0.0)
    if is_complex:
        coef = coef * np.exp(1j * float(np.float32(wavenum)) * tau)
    g = np.asarray(grad_out).reshape(B, N)
    gx = np.zeros((B, C, T, S), dtype=np.complex128 if is_complex else np.float64)
    ch = np.broadcast_to(np.arange(C)[None, :, None], k.shape)
    seqc = np.where(valid, seq, 0)
    for b in range(B):
        v = np.conj(coef) * g[b][None, None, :]
        np.add.at(gx[b], (ch, seqc, k), v * (1 - t))
        np.add.at(gx[b], (ch, seqc, np.minimum(k + 1, S - 1)), np.where(valid, v * t, 0))
    return gx if batched else gx[0]
