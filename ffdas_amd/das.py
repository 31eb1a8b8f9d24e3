"""``das`` / ``das_sparse``: the reference's Python operators on the gfx950 library.

Same signature, keyword-only arguments, validation errors, casts and output conventions as
the reference (bindings/python/ffdas/das.py:64-287); the C call goes through the C-ABI
(libffdas.so) with torch's current stream.  Additions, all keyword-only and defaulting to the
reference behaviour:

* ``scale``: global delay scale s, tau = s * (offsets + r) (differentiable sound-speed scale);
* ``use_bf16``: store the channel data as bfloat16 (like ``use_fp16``);
* autograd: when x, offsets, weights or a tensor ``scale`` require grad, the call goes
  through :class:`ffdas_amd.autograd.DasFunction`, whose backward runs the library's
  hand-written adjoint kernels.
"""
from __future__ import annotations

import ctypes
import warnings

import torch

from . import _core
from ._core import TensorView, as_tensor, check_device_tensor, element_code, handle_for
from ._lib import Algorithm, ComputeType, DasProblem, DasTuning, DType, TensorDesc, check, lib


class DowncastWarning(UserWarning):
    """An input was cast to the library's argument type (reference _core/tensor.py:12-42)."""


def _astype(t, dtype: torch.dtype, name: str):
    if t is None:
        return None
    t = as_tensor(t)
    if t.dtype != dtype:
        warnings.warn(f"input of type {t.dtype} will be downcast to {dtype}", DowncastWarning,
                      stacklevel=4)
        t = t.to(dtype)
    return t.contiguous()


def _compute_type(x, use_fp16: bool, use_bf16: bool) -> ComputeType:
    """Reference das.py:20-30 (plus the bf16 extension)."""
    if use_fp16:
        return ComputeType.FP16
    if use_bf16:
        return ComputeType.BF16
    code = element_code(x)
    if code in (DType.R_64F, DType.C_64F):
        return ComputeType.FP64
    return ComputeType.FP32


def _out_dtype(code: int) -> torch.dtype:
    if code == DType.R_64F:
        return torch.float64
    if code == DType.C_64F:
        return torch.complex128
    if code in (DType.C_16F, DType.C_32F, DType.C_16BF):
        return torch.complex64
    return torch.float32


def _out_code(dtype: torch.dtype) -> int:
    return {torch.float32: DType.R_32F, torch.float64: DType.R_64F,
            torch.complex64: DType.C_32F, torch.complex128: DType.C_64F}[dtype]


def _x_desc(x, channels_trailing: bool) -> TensorDesc:
    """Descriptor of x; channels_trailing swaps the channel and sequence dimensions
    (reference _ffdas.cpp:507-519)."""
    shape, strides = _core.shape_strides(x)
    dims, st = list(shape), list(strides)
    if channels_trailing:
        c = 1 if len(dims) == 4 else 0
        dims[c], dims[c + 1] = dims[c + 1], dims[c]
        st[c], st[c + 1] = st[c + 1], st[c]
    return TensorDesc(dims, st, element_code(x))


def _validate_common(x, dstpos, offsets, weights):
    """Reference das.py:107-123."""
    if dstpos.ndim < 2 or dstpos.shape[-1] != 3:
        raise ValueError("dstpos must have dimensions (..., 3)")
    spatial = tuple(dstpos.shape[:-1])
    out_dims = spatial
    if x.ndim == 4:
        out_dims = (x.shape[0], *out_dims)
    if tuple(offsets.shape[1:]) != spatial:
        raise ValueError(f"offsets shape {tuple(offsets.shape)} must match target layout {spatial}")
    if tuple(weights.shape[1:]) != spatial:
        raise ValueError(f"weights shape {tuple(weights.shape)} must match target layout {spatial}")
    return spatial, out_dims


def _beta_buffer(out_dtype: torch.dtype):
    if out_dtype == torch.float64:
        return (ctypes.c_double * 1)(0.0)
    if out_dtype == torch.complex128:
        return (ctypes.c_double * 2)(0.0, 0.0)
    if out_dtype == torch.complex64:
        return (ctypes.c_float * 2)(0.0, 0.0)
    return (ctypes.c_float * 1)(0.0)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _prepare(x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices):
    x = as_tensor(x)
    srcpos, dstpos, offsets, weights = (as_tensor(a) for a in (srcpos, dstpos, offsets, weights))
    if sparse_indices is not None:
        sparse_indices = as_tensor(sparse_indices)
    spatial, out_dims = _validate_common(x, dstpos, offsets, weights)
    return x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices, spatial, out_dims


def _forward_call(x, srcpos, dstpos, offsets, weights, srcdir, wavenum, sparse_indices,
                  scale, algorithm, compute_type, channels_trailing, out, out_dims):
    """Cast the geometry, allocate out, and call ffdas_das / ffdas_das_sparse / ffdas_das_ex."""
    check_device_tensor(x, "x")
    dev = x.device
    srcpos = _astype(srcpos, torch.float32, "srcpos")
    dstpos = _astype(dstpos, torch.float32, "dstpos")
    offsets = _astype(offsets, torch.float32, "offsets")
    weights = _astype(weights, torch.float32, "weights")
    srcdir = _astype(srcdir, torch.float32, "srcdir")
    if sparse_indices is not None:
        sparse_indices = _astype(sparse_indices, torch.int32, "sparse_indices")
    for name, t in (("srcpos", srcpos), ("dstpos", dstpos), ("offsets", offsets),
                    ("weights", weights), ("srcdir", srcdir), ("sparse_indices", sparse_indices)):
        if t is not None:
            check_device_tensor(t, name, dev)
    odt = _out_dtype(element_code(x))
    if out is None:
        out = torch.empty(out_dims, dtype=odt, device=dev)
    else:
        check_device_tensor(out, "out", dev)
    x_desc = _x_desc(x, channels_trailing)
    o_desc = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
    beta = _beta_buffer(out.dtype)
    L = lib()
    with torch.cuda.device(dev):
        h = handle_for(x)
        if scale != 1.0:
            K = 0 if sparse_indices is None else int(sparse_indices.shape[0])
            rc = L.ffdas_das_ex(h.h, _ptr(srcpos), _ptr(srcdir), float(wavenum), float(scale),
                                x_desc.ptr, _ptr(x), _ptr(dstpos), _ptr(offsets), _ptr(weights),
                                K, _ptr(sparse_indices), ctypes.cast(beta, ctypes.c_void_p),
                                o_desc.ptr, _ptr(out), int(compute_type), int(algorithm))
            check(rc, "ffdas_das_ex")
        elif sparse_indices is None:
            rc = L.ffdas_das(h.h, _ptr(srcpos), _ptr(srcdir), float(wavenum), x_desc.ptr, _ptr(x),
                             _ptr(dstpos), _ptr(offsets), _ptr(weights),
                             ctypes.cast(beta, ctypes.c_void_p), o_desc.ptr, _ptr(out),
                             int(compute_type), int(algorithm))
            check(rc, "ffdas_das")
        else:
            rc = L.ffdas_das_sparse(h.h, _ptr(srcpos), _ptr(srcdir), float(wavenum), x_desc.ptr,
                                    _ptr(x), _ptr(dstpos), _ptr(offsets), _ptr(weights),
                                    int(sparse_indices.shape[0]), _ptr(sparse_indices),
                                    ctypes.cast(beta, ctypes.c_void_p), o_desc.ptr, _ptr(out),
                                    int(compute_type), int(algorithm))
            check(rc, "ffdas_das_sparse")
    return out


def _needs_grad(*ts) -> bool:
    if not torch.is_grad_enabled():
        return False
    return any(isinstance(t, torch.Tensor) and t.requires_grad for t in ts)


def das(x, srcpos, dstpos, offsets, weights, *, srcdir=None, wavenum: float = 0.0,
        algorithm: Algorithm = Algorithm.DEFAULT, use_fp16: bool = False,
        channels_trailing: bool = False, out=None, scale=1.0, use_bf16: bool = False):
    """Delay-and-sum beamforming, compounding over the sequence (transmit) dimension.

    For each target, computes a weighted sum of linearly interpolated samples from all
    channels and sequence events.  Positions are in sampling wavelengths (c / fs), offsets
    in samples (reference das.py:78-106, docs/api/das.md).

    Args:
        x: channel data ``([batch,] channels, seq, samples)`` (torch ROCm tensor, DLPack
            array or :class:`TensorView`).
        srcpos: channel positions ``(channels, 3)``.
        dstpos: target positions ``(..., 3)``; the spatial dims are ``dstpos.shape[:-1]``.
        offsets: per-target transmit delays in samples ``(seq, ...)``.
        weights: per-target apodization ``(seq, ...)``.
        srcdir: directivity ``(channels, 4)``: unit normal and cosine of the half-angle.
        wavenum: phase rotation in rad/sample for complex data (0 disables).
        algorithm: accepted for compatibility; every value but ALG3 runs the same kernels.
        use_fp16 / use_bf16: store the channel data in half / bfloat16 (FP32 arithmetic).
        channels_trailing: x is ``([batch,] seq, channels, samples)``.
        out: preallocated output.
        scale: delay scale s (tau = s * (offsets + r)); float or 0-d tensor.

    Returns:
        ``([batch,] ...)``; float32 for real 16/32-bit data, complex64 for complex, float64 /
        complex128 for double data.
    """
    x, srcpos, dstpos, offsets, weights, srcdir, _, spatial, out_dims = _prepare(
        x, srcpos, dstpos, offsets, weights, srcdir, None)
    ct = _compute_type(x, use_fp16, use_bf16)
    if _needs_grad(x, offsets, weights, scale) and not isinstance(x, TensorView):
        from .autograd import DasFunction
        return DasFunction.apply(x, offsets, weights, scale, srcpos, dstpos, srcdir, None,
                                 float(wavenum), ct, bool(channels_trailing), algorithm)
    s = float(scale) if not isinstance(scale, torch.Tensor) else float(scale.detach().item())
    return _forward_call(x, srcpos, dstpos, offsets, weights, srcdir, wavenum, None, s,
                         algorithm, ct, channels_trailing, out, out_dims)


def das_sparse(x, srcpos, dstpos, offsets, weights, sparse_indices, *, srcdir=None,
               wavenum: float = 0.0, algorithm: Algorithm = Algorithm.DEFAULT,
               use_fp16: bool = False, channels_trailing: bool = False, out=None, scale=1.0,
               use_bf16: bool = False):
    """Sparse compounding delay-and-sum: each target compounds the n sequence events listed
    in ``sparse_indices`` ``(n, ...)`` (int32, 0-based); ``offsets`` and ``weights`` are
    ``(n, ...)`` (reference das.py:188-287).  Other arguments as :func:`das`."""
    x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices, spatial, out_dims = _prepare(
        x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices)
    if tuple(sparse_indices.shape[1:]) != spatial:
        raise ValueError(f"sparse_indices shape {tuple(sparse_indices.shape)} must match target "
                         f"layout {spatial}")
    sparse_count = sparse_indices.shape[0]
    if offsets.shape[0] != sparse_count or weights.shape[0] != sparse_count:
        raise ValueError(f"offsets and weights leading dimension must match sparse count "
                         f"{sparse_count}")
    ct = _compute_type(x, use_fp16, use_bf16)
    if _needs_grad(x, offsets, weights, scale) and not isinstance(x, TensorView):
        from .autograd import DasFunction
        return DasFunction.apply(x, offsets, weights, scale, srcpos, dstpos, srcdir,
                                 sparse_indices, float(wavenum), ct, bool(channels_trailing),
                                 algorithm)
    s = float(scale) if not isinstance(scale, torch.Tensor) else float(scale.detach().item())
    return _forward_call(x, srcpos, dstpos, offsets, weights, srcdir, wavenum, sparse_indices, s,
                         algorithm, ct, channels_trailing, out, out_dims)


def das_backward(x, srcpos, dstpos, offsets, weights, grad_out, *, srcdir=None, wavenum=0.0,
                 sparse_indices=None, scale=1.0, use_fp16=False, use_bf16=False,
                 channels_trailing=False, need_x=True, need_offsets=True, need_weights=True,
                 need_scale=True, algorithm: Algorithm = Algorithm.DEFAULT, reuse_x=False):
    """Adjoint of :func:`das` / :func:`das_sparse` via ffdas_das_backward.

    Returns ``(grad_x, grad_offsets, grad_weights, grad_scale)``; unrequested entries are None.
    grad_x has x's shape (float32/complex64, float64/complex128 for double data).
    ``reuse_x``: x is unchanged since the last forward on this device, so the converted copy
    that forward made may be gathered from again (ffdas_set_reuse_x)."""
    x = as_tensor(x)
    check_device_tensor(x, "x")
    dev = x.device
    srcpos = _astype(srcpos, torch.float32, "srcpos")
    dstpos = _astype(dstpos, torch.float32, "dstpos")
    offsets = _astype(offsets, torch.float32, "offsets")
    weights = _astype(weights, torch.float32, "weights")
    srcdir = _astype(srcdir, torch.float32, "srcdir")
    if sparse_indices is not None:
        sparse_indices = _astype(sparse_indices, torch.int32, "sparse_indices")
    ct = _compute_type(x, use_fp16, use_bf16)
    xc = element_code(x)
    gdt = _out_dtype(xc)
    grad_out = as_tensor(grad_out).to(gdt).contiguous()
    x_desc = _x_desc(x, channels_trailing)
    g_desc = TensorDesc(tuple(grad_out.shape), tuple(grad_out.stride()), _out_code(gdt))
    shape, _ = _core.shape_strides(x)
    dims = list(shape)
    if channels_trailing:
        c = 1 if len(dims) == 4 else 0
        dims[c], dims[c + 1] = dims[c + 1], dims[c]
    gx = torch.empty(dims, dtype=gdt, device=dev) if need_x else None
    gx_desc = TensorDesc(dims, None, _out_code(gdt)) if need_x else None
    K = offsets.shape[0]
    go = torch.empty_like(offsets) if need_offsets else None
    gw = torch.empty_like(weights) if need_weights else None
    gs = torch.empty((), dtype=torch.float32, device=dev) if need_scale else None
    L = lib()
    with torch.cuda.device(dev):
        h = handle_for(x)
        check(L.ffdas_set_reuse_x(h.h, int(bool(reuse_x))), "ffdas_set_reuse_x")
        rc = L.ffdas_das_backward(
            h.h, _ptr(srcpos), _ptr(srcdir), float(wavenum), float(scale), x_desc.ptr, _ptr(x),
            _ptr(dstpos), _ptr(offsets), _ptr(weights),
            int(sparse_indices.shape[0]) if sparse_indices is not None else 0,
            _ptr(sparse_indices), g_desc.ptr, _ptr(grad_out),
            gx_desc.ptr if gx_desc is not None else None, _ptr(gx), _ptr(go), _ptr(gw), _ptr(gs),
            int(ct), int(algorithm))
        L.ffdas_set_reuse_x(h.h, 0)
        check(rc, "ffdas_das_backward")
    if gx is not None and channels_trailing:
        c = 1 if len(dims) == 4 else 0
        gx = gx.transpose(c, c + 1)
    return gx, go, gw, gs


def autotune(x, srcpos, dstpos, offsets, weights, *, srcdir=None, wavenum=0.0,
             sparse_indices=None, scale=1.0, use_fp16=False, use_bf16=False,
             channels_trailing=False, adjoint=True, reps=5,
             algorithm: Algorithm = Algorithm.DEFAULT) -> dict:
    """Time candidate launch parameters for this problem (ffdas_das_autotune) and keep the
    fastest in the library's tuning cache: later :func:`das` / :func:`das_backward` calls with
    the same signature (device, dtypes, channels, transmits, samples, frames, grid, sparse
    count, directivity) use them.  With ``adjoint`` the grad_x and parameter kernels are tuned
    too.  Synchronous.  Returns the chosen parameters and the measured times [ms]."""
    x = as_tensor(x)
    check_device_tensor(x, "x")
    dev = x.device
    srcpos = _astype(srcpos, torch.float32, "srcpos")
    dstpos = _astype(dstpos, torch.float32, "dstpos")
    offsets = _astype(offsets, torch.float32, "offsets")
    weights = _astype(weights, torch.float32, "weights")
    srcdir = _astype(srcdir, torch.float32, "srcdir")
    if sparse_indices is not None:
        sparse_indices = _astype(sparse_indices, torch.int32, "sparse_indices")
    ct = _compute_type(x, use_fp16, use_bf16)
    odt = _out_dtype(element_code(x))
    shape, _ = _core.shape_strides(x)
    batch = (shape[0],) if len(shape) == 4 else ()
    out = torch.empty(batch + tuple(dstpos.shape[:-1]), dtype=odt, device=dev)
    o_desc = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
    x_desc = _x_desc(x, channels_trailing)
    dims = list(shape)
    if channels_trailing:
        c = 1 if len(dims) == 4 else 0
        dims[c], dims[c + 1] = dims[c + 1], dims[c]
    gx = torch.empty(dims, dtype=odt, device=dev) if adjoint else None
    gx_desc = TensorDesc(dims, None, _out_code(odt)) if adjoint else None
    res = DasTuning()
    L = lib()
    with torch.cuda.device(dev):
        h = handle_for(x)
        rc = L.ffdas_das_autotune(
            h.h, _ptr(srcpos), _ptr(srcdir), float(wavenum), float(scale), x_desc.ptr, _ptr(x),
            _ptr(dstpos), _ptr(offsets), _ptr(weights),
            int(sparse_indices.shape[0]) if sparse_indices is not None else 0,
            _ptr(sparse_indices), o_desc.ptr, _ptr(out),
            gx_desc.ptr if gx_desc is not None else None, _ptr(gx), int(ct), int(algorithm),
            int(reps), ctypes.byref(res))
        check(rc, "ffdas_das_autotune")
    return res.as_dict()


def save_tuning(path) -> None:
    """Write the tuning cache to a text file (one line per problem signature)."""
    check(lib().ffdas_tune_save(str(path).encode()), "ffdas_tune_save")


def load_tuning(path) -> None:
    """Merge a tuning file written by :func:`save_tuning` into the cache."""
    check(lib().ffdas_tune_load(str(path).encode()), "ffdas_tune_load")


def clear_tuning() -> None:
    check(lib().ffdas_tune_clear(), "ffdas_tune_clear")


def tuning_entries() -> int:
    return int(lib().ffdas_tune_count())


def das_batched(problems, *, use_fp16=False, use_bf16=False,
                algorithm: Algorithm = Algorithm.DEFAULT):
    """Several forward problems in one call (ffdas_das_batched): mixed probes, voxel grids,
    frame ensembles and sparse tables; problems that share storage type and frame chunking
    run in one gather launch.  Each problem is a dict with the arguments of :func:`das` /
    :func:`das_sparse` -- ``x, srcpos, dstpos, offsets, weights`` and optionally
    ``srcdir, wavenum, sparse_indices, scale, channels_trailing, out`` -- and gets the same
    result.  Returns the list of outputs."""
    keep, structs, outs = [], [], []
    dev = None
    for q in problems:
        x = as_tensor(q["x"])
        check_device_tensor(x, "x", dev)
        dev = x.device
        srcpos = _astype(q["srcpos"], torch.float32, "srcpos")
        dstpos = _astype(q["dstpos"], torch.float32, "dstpos")
        offsets = _astype(q["offsets"], torch.float32, "offsets")
        weights = _astype(q["weights"], torch.float32, "weights")
        srcdir = _astype(q.get("srcdir"), torch.float32, "srcdir")
        si = q.get("sparse_indices")
        if si is not None:
            si = _astype(si, torch.int32, "sparse_indices")
        _validate_common(x, dstpos, offsets, weights)
        shape, _ = _core.shape_strides(x)
        batch = (shape[0],) if len(shape) == 4 else ()
        odt = _out_dtype(element_code(x))
        out = q.get("out")
        if out is None:
            out = torch.empty(batch + tuple(dstpos.shape[:-1]), dtype=odt, device=dev)
        x_desc = _x_desc(x, bool(q.get("channels_trailing", False)))
        o_desc = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
        beta = _beta_buffer(out.dtype)
        keep += [x, srcpos, dstpos, offsets, weights, srcdir, si, x_desc, o_desc, beta]
        structs.append(DasProblem(
            _ptr(srcpos), _ptr(srcdir), float(q.get("wavenum", 0.0)), float(q.get("scale", 1.0)),
            x_desc.ptr, _ptr(x), _ptr(dstpos), _ptr(offsets), _ptr(weights),
            int(si.shape[0]) if si is not None else 0, _ptr(si),
            ctypes.cast(beta, ctypes.c_void_p), o_desc.ptr, _ptr(out)))
        outs.append(out)
    if not structs:
        return []
    arr = (DasProblem * len(structs))(*structs)
    ct = _compute_type(as_tensor(problems[0]["x"]), use_fp16, use_bf16)
    with torch.cuda.device(dev):
        h = handle_for(as_tensor(problems[0]["x"]))
        check(lib().ffdas_das_batched(h.h, len(structs), arr, int(ct), int(algorithm)),
              "ffdas_das_batched")
    return outs


class CapturedCall:
    """A library call sequence captured once into a HIP graph (torch.cuda.CUDAGraph) and
    replayed: ``fn`` is run once eagerly (it sizes the handle's workspace and resolves the
    tuning), then captured; ``replay()`` re-runs every captured launch on the current stream
    with the captured pointers, so inputs are updated in place and outputs read from the
    tensors ``fn`` returned at capture time."""

    def __init__(self, fn):
        fn()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = fn()
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()
        return self.outputs
