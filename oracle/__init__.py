"""CPU oracle for ffdas `das` / `das_sparse` and their adjoint (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py import this
package, and only as the checker.  The product (``ffdas_amd`` + ``libffdas.so``) never
imports it.

Two independent restatements of the reference's arithmetic definition
(src/das/das_alg1_kernels.cuh:41-106, docs/api/das.md:5,37-48,90-100):

* ``das`` / ``das_adjoint`` -> ``libffdas_oracle.so`` (das_oracle.cpp), multi-threaded,
  FP64 (the correctness oracle) or FP32 (the CPU baseline);
* ``brute.py`` -> a numpy FP64 restatement written separately, used to cross-check the C++
  oracle on small inputs and to write the golden fixtures.

Pinning (what checks each function of the oracle against something it does not restate):
  forward  -- reference known answers (bindings/python/tests/test_das.py:83-132), closed
              forms (ramp, constant, complex constant, two-tap interpolation), invariants
              (linearity, additivity over transmits, sparse == dense, top-K sparse == dense),
              brute force in numpy, golden fixtures;
  adjoint  -- the dot-product test <x, A^T y> = <A x, y> against the pinned forward, and
              central finite differences of the pinned forward for offsets / weights /
              scale.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libffdas_oracle.so"
_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with make (g++)."""
    import subprocess

    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-C", str(_HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        f = ctypes.c_float
        p = ctypes.c_void_p
        i = ctypes.c_int
        d = ctypes.c_double
        lib.ffdas_oracle_das.argtypes = [i, i, i, i, i, i, i, i, i, p, p, f, d, p, p, p, p, p, d,
                                         d, p]
        lib.ffdas_oracle_das.restype = i
        lib.ffdas_oracle_das_adjoint.argtypes = [i, i, i, i, i, i, i, i, i, p, p, f, d, p, p, p,
                                                 p, p, p, p, p, p, p]
        lib.ffdas_oracle_das_adjoint.restype = i
        lib.ffdas_oracle_hardware_threads.restype = i
        _lib = lib
    return _lib


def hardware_threads() -> int:
    return int(_load().ffdas_oracle_hardware_threads())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _prep(x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices, bits, channels_trailing):
    real = np.float64 if bits == 64 else np.float32
    x = np.asarray(x)
    is_complex = np.iscomplexobj(x)
    if x.ndim == 3:
        x = x[None]
        batched = False
    elif x.ndim == 4:
        batched = True
    else:
        raise ValueError("x must be 3-D or 4-D")
    if channels_trailing:
        x = np.swapaxes(x, 1, 2)
    B, C, T, S = x.shape
    if is_complex:
        xr = np.ascontiguousarray(x.astype(np.complex128 if bits == 64 else np.complex64))
        xr = xr.view(real)
    else:
        xr = np.ascontiguousarray(x, dtype=real)
    dstpos = np.ascontiguousarray(dstpos, dtype=np.float32)
    spatial = dstpos.shape[:-1]
    N = int(np.prod(spatial))
    offsets = np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1, N)
    weights = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1, N)
    K = offsets.shape[0]
    if sparse_indices is not None:
        sparse_indices = np.ascontiguousarray(sparse_indices, dtype=np.int32).reshape(-1, N)
    srcpos = np.ascontiguousarray(srcpos, dtype=np.float32)
    if srcdir is not None:
        srcdir = np.ascontiguousarray(srcdir, dtype=np.float32)
    return (xr, is_complex, batched, (B, C, T, S, N, K), spatial, srcpos, dstpos, offsets,
            weights, srcdir, sparse_indices, real)


def das(x, srcpos, dstpos, offsets, weights, *, srcdir=None, wavenum=0.0, sparse_indices=None,
        scale=1.0, bits=64, threads=None, channels_trailing=False, beta=0.0, out=None):
    """Forward DAS on the CPU.  Returns ([B,] *spatial) in float64/complex128 (bits=64) or
    float32/complex64 (bits=32)."""
    lib = _load()
    (xr, cplx, batched, (B, C, T, S, N, K), spatial, srcpos, dstpos, offsets, weights, srcdir,
     sparse_indices, real) = _prep(x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices,
                                   bits, channels_trailing)
    cw = 2 if cplx else 1
    if out is None:
        o = np.zeros((B, N, cw), dtype=real)
    else:
        o = np.ascontiguousarray(np.asarray(out).reshape(B, N))
        o = (o.astype(np.complex128 if bits == 64 else np.complex64).view(real).reshape(B, N, 2)
             if cplx else o.astype(real).reshape(B, N, 1)).copy()
    beta = complex(beta)
    rc = lib.ffdas_oracle_das(bits, int(threads or hardware_threads()), B, C, T, S, N, K,
                              int(cplx), _ptr(srcpos), _ptr(srcdir), float(wavenum), float(scale),
                              _ptr(xr), _ptr(dstpos), _ptr(offsets), _ptr(weights),
                              _ptr(sparse_indices), beta.real, beta.imag, _ptr(o))
    if rc != 0:
        raise ValueError(f"oracle rejected arguments (code {rc})")
    res = o.reshape(B, N, cw)
    res = (res[..., 0] + 1j * res[..., 1]) if cplx else res[..., 0]
    res = res.reshape((B,) + tuple(spatial))
    return res if batched else res[0]


def das_adjoint(x, srcpos, dstpos, offsets, weights, grad_out, *, srcdir=None, wavenum=0.0,
                sparse_indices=None, scale=1.0, bits=64, threads=None):
    """Adjoint of ``das`` on the CPU.  Returns dict with grad_x (shape of x), grad_offsets,
    grad_weights (shape of offsets) and grad_scale (float)."""
    lib = _load()
    x_shape = np.shape(x)
    (xr, cplx, batched, (B, C, T, S, N, K), spatial, srcpos, dstpos, offsets_c, weights_c,
     srcdir, sparse_indices, real) = _prep(x, srcpos, dstpos, offsets, weights, srcdir,
                                           sparse_indices, bits, False)
    cw = 2 if cplx else 1
    g = np.asarray(grad_out).reshape(B, N)
    if cplx:
        gr = np.ascontiguousarray(g.astype(np.complex128 if bits == 64 else np.complex64)).view(real)
    else:
        gr = np.ascontiguousarray(g, dtype=real)
    gx = np.zeros((B, C, T, S * cw), dtype=real)
    go = np.zeros((K, N), dtype=real)
    gw = np.zeros((K, N), dtype=real)
    gs = np.zeros(1, dtype=real)
    rc = lib.ffdas_oracle_das_adjoint(bits, int(threads or hardware_threads()), B, C, T, S, N, K,
                                      int(cplx), _ptr(srcpos), _ptr(srcdir), float(wavenum),
                                      float(scale), _ptr(xr), _ptr(dstpos), _ptr(offsets_c),
                                      _ptr(weights_c), _ptr(sparse_indices), _ptr(gr), _ptr(gx),
                                      _ptr(go), _ptr(gw), _ptr(gs))
    if rc != 0:
        raise ValueError(f"oracle rejected arguments (code {rc})")
    if cplx:
        gx = gx.reshape(B, C, T, S, 2)
        gx = gx[..., 0] + 1j * gx[..., 1]
    gx = gx.reshape(x_shape)
    kshape = np.shape(offsets)
    return {"grad_x": gx, "grad_offsets": go.reshape(kshape), "grad_weights": gw.reshape(kshape),
            "grad_scale": float(gs[0])}


def rel_l2(a, b) -> float:
    """||a - b|| / ||b|| in float64."""
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b)
                   else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
