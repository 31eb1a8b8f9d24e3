"""The MATLAB MEX gateways' argument handling, restated over the C-ABI (no MATLAB here).

A MATLAB gpuArray of size (d1, ..., dn) is column-major; the reference's gateways view it
as the row-major array of dims (dn, ..., d1) with C-contiguous strides
(bindings/matlab/+ffdas/+core/src/_mex_ndarray.h:300-311), and MATLAB drops trailing
singleton dimensions (an array keeps at least two).  Here such an array is a torch tensor
holding that row-major view; `matlab_squeeze` applies MATLAB's rule to it.

`mex_das` / `mex_das_sparse` follow ffdas_das.cpp:13-151 and ffdas_das_sparse.cpp:13-148:
the dimension checks and their messages, the re-insertion of a squeezed (batch, sequence)
pair, the (channels, sequence) swap for channels_trailing=false done as a strided
descriptor (no copy), the output allocation, the host zero `beta`, the compute type, and
the 1-based to 0-based index shift of das_sparse.m:52.  They call `ffdas_das` /
`ffdas_das_sparse` through ctypes exactly as the gateways call them from C++.

Two gateway quirks are not reproduced: after the channels_trailing=false permute the
gateways read the sequence length from the channel axis (ffdas_das.cpp:59), and the sparse
gateway passes sparse_indices.shape(0) as the sparse count even when MATLAB has squeezed a
count of 1 away (ffdas_das_sparse.cpp:131).  Here the sequence length is the transmit
axis and a squeezed index table means a count of 1.
"""
from __future__ import annotations

import ctypes

import torch

from ffdas_amd._core import element_code, handle_for
from ffdas_amd._lib import ComputeType, TensorDesc, check, lib


class MexError(RuntimeError):
    """mexErrMsgIdAndTxt."""


def matlab_squeeze(t: torch.Tensor) -> torch.Tensor:
    """Drop MATLAB-trailing (row-major leading) singleton dims, keeping at least two."""
    while t.dim() > 2 and t.shape[0] == 1:
        t = t[0]
    return t


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _x_view(x: torch.Tensor, channels_trailing: bool, name: str) -> torch.Tensor:
    # ffdas_das.cpp:39-57
    if x.dim() < 3 and not (not channels_trailing and x.dim() == 2):
        raise MexError(f"{name}: x must have 3 or 4 dimensions (got {x.dim()})")
    if x.dim() > 4:
        raise MexError(f"{name}: x must have 3 or 4 dimensions (got {x.dim()})")
    if not channels_trailing:
        if x.dim() == 2:            # batch and sequence both squeezed away by MATLAB
            x = x.reshape(1, x.shape[0], x.shape[1])
        x = x.permute(1, 0, 2) if x.dim() == 3 else x.permute(0, 2, 1, 3)
    return x


def _check_geometry(name, x, srcpos, dstpos, srcdir, tables):
    sequence = x.shape[-2]          # x is ([B,] C, T, S) here
    if srcpos.dim() != 2 or srcpos.shape[1] != 3:
        raise MexError(f"{name}: srcpos must have shape (3, channels)")
    if srcdir is not None and srcdir.numel() > 0:
        if srcdir.dtype != torch.float32 or srcdir.dim() != 2 or srcdir.shape[1] != 4:
            raise MexError(f"{name}: srcdir must be a single array of shape (4, channels)")
    dstnd = dstpos.dim()
    if dstnd < 2 or dstpos.shape[-1] != 3:
        raise MexError(f"{name}: dstpos must have shape (3, ...)")
    first = tables[0]
    if any(t.shape != first.shape for t in tables[1:]):
        raise MexError(f"{name}: offsets and weights must have the same shape")
    return sequence, dstnd, first


def _check_tables(name, sequence, dstnd, dstpos, t, lead):
    # ffdas_das.cpp:75-92 (lead = sequence for das, the sparse count for das_sparse)
    if lead == 1 and t.dim() != dstnd - 1:
        raise MexError(f"{name}: offsets and weights must have the same number of dimensions "
                       f"as dstpos")
    if lead > 1 and t.dim() - 1 != dstnd - 1:
        raise MexError(f"{name}: offsets and weights must have the same number of dimensions "
                       f"as dstpos")
    if lead > 1 and t.shape[0] != lead:
        raise MexError(f"{name}: leading dimension of offsets/weights must match the channel "
                       f"dimension of x")
    for i in range(dstnd - 1):
        if t.shape[i + 1 if lead > 1 else i] != dstpos.shape[i]:
            raise MexError(f"{name}: spatial dimensions of offsets/weights must match dstpos")


def _out_dtype(x):
    return {torch.float16: torch.float32, torch.bfloat16: torch.float32,
            torch.complex32: torch.complex64}.get(x.dtype, x.dtype)


def _compute_type(x, use_fp16):
    if x.dtype in (torch.float64, torch.complex128):
        return ComputeType.FP64
    return ComputeType.FP16 if use_fp16 else ComputeType.FP32


def _call(name, x, srcpos, dstpos, offsets, weights, srcdir, wavenum, algorithm, use_fp16,
          sparse=None, count=0):
    have_batch = x.dim() == 4
    spatial = tuple(dstpos.shape[:-1])
    out_dims = ((x.shape[0],) if have_batch else ()) + spatial
    out = torch.empty(out_dims, dtype=_out_dtype(x), device=x.device)
    x_desc = TensorDesc(tuple(x.shape), tuple(x.stride()), element_code(x))
    o_desc = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
    beta = ctypes.create_string_buffer(16)         # calloc(1, type size): zero
    sd = srcdir if srcdir is not None and srcdir.numel() > 0 else None
    L = lib()
    with torch.cuda.device(x.device):
        h = handle_for(x)
        if sparse is None:
            rc = L.ffdas_das(h.h, _ptr(srcpos), _ptr(sd), float(wavenum), x_desc.ptr, _ptr(x),
                             _ptr(dstpos), _ptr(offsets), _ptr(weights),
                             ctypes.cast(beta, ctypes.c_void_p), o_desc.ptr, _ptr(out),
                             int(_compute_type(x, use_fp16)), int(algorithm))
        else:
            rc = L.ffdas_das_sparse(h.h, _ptr(srcpos), _ptr(sd), float(wavenum), x_desc.ptr,
                                    _ptr(x), _ptr(dstpos), _ptr(offsets), _ptr(weights),
                                    int(count), _ptr(sparse), ctypes.cast(beta, ctypes.c_void_p),
                                    o_desc.ptr, _ptr(out), int(_compute_type(x, use_fp16)),
                                    int(algorithm))
    try:
        check(rc, name)
    except RuntimeError as e:
        raise MexError(f"{name}: returned error {rc}: {e}") from e
    return out


def mex_das(x, srcpos, dstpos, offsets, weights, srcdir=None, wavenum=0.0, algorithm=0,
            use_fp16=False, channels_trailing=True):
    """ffdas.das from MATLAB: ffdas_das.cpp:13-151 (das.m supplies the defaults)."""
    name = "ffdas_das"
    x = _x_view(x, channels_trailing, name)
    sequence, dstnd, t = _check_geometry(name, x, srcpos, dstpos, srcdir, [offsets, weights])
    _check_tables(name, sequence, dstnd, dstpos, t, sequence)
    return _call(name, x, srcpos, dstpos, offsets, weights, srcdir, wavenum, algorithm, use_fp16)


def mex_das_sparse(x, srcpos, dstpos, offsets, weights, sparse_indices_1based, srcdir=None,
                   wavenum=0.0, algorithm=0, use_fp16=False, channels_trailing=True):
    """ffdas.das_sparse from MATLAB: das_sparse.m:45-53 (1-based indices, minus 1) and
    ffdas_das_sparse.cpp:13-148."""
    name = "ffdas_das_sparse"
    sparse = (sparse_indices_1based.to(torch.int32) - 1).contiguous()
    x = _x_view(x, channels_trailing, name)
    _, dstnd, t = _check_geometry(name, x, srcpos, dstpos, srcdir, [offsets, weights, sparse])
    count = t.shape[0] if t.dim() == dstnd else 1
    _check_tables(name, count, dstnd, dstpos, t, count)
    return _call(name, x, srcpos, dstpos, offsets, weights, srcdir, wavenum, algorithm, use_fp16,
                 sparse=sparse, count=count)


def mex_truncate_rank(x: torch.Tensor, start_1based: int, stop: int | None = None):
    """truncate_rank.m + ffdas_truncate_rank.cpp:13-70: MATLAB's last dimension (the leading
    dimension of the row-major view) is kept, the others are flattened; start is 1-based and
    stop inclusive (so the same number as the 0-based exclusive stop); default stop
    min(m, n)."""
    name = "ffdas_truncate_rank"
    if x.dim() < 2:
        raise MexError(f"{name}:error: input must have at least two dimensions (got {x.dim()})")
    n = int(x.shape[0])
    m = 1
    for d in x.shape[1:]:
        m *= int(d)
    if stop is None:                       # truncate_rank.m: stop = min(m, n)
        stop = min(m, n)
    start = int(start_1based) - 1
    if start < 0 or start >= min(m, n):
        raise MexError(f"{name}:error: start must be > 0 and < min(m, n) (got {start})")
    if stop <= start or stop > min(m, n):
        raise MexError(f"{name}:error: stop must be > start and <= min(m, n) (got {stop})")
    xv = x.reshape(n, m)
    out = torch.empty_like(xv)
    x_desc = TensorDesc((n, m), tuple(xv.stride()), element_code(xv))
    o_desc = TensorDesc((n, m), tuple(out.stride()), element_code(out))
    with torch.cuda.device(x.device):
        h = handle_for(x)
        check(lib().ffdas_truncate_rank(h.h, x_desc.ptr, xv.data_ptr(), start, int(stop),
                                        o_desc.ptr, out.data_ptr()), name)
    return out.reshape(x.shape)


def mex_interpolate(gridpos, x, querypos, mode: str, fill_value):
    """interpolate.m + ffdas_interpolate.cpp:12-120: gridpos (3, nx, ny, nz) in MATLAB (the
    row-major view (nz, ny, nx, 3)), x (nx, ny, nz[, batch...]), querypos (3, ...), a mode
    string and a one-element host fill value; a plan is created, evaluated and destroyed per
    call; the result has x's leading (MATLAB-trailing) dims followed by the query dims."""
    name = "ffdas_interpolate"
    if gridpos.dim() != 4 or gridpos.shape[3] != 3:
        raise MexError(f"{name}:error: gridpos must have shape (3, nx, ny, nz)")
    fill = torch.as_tensor(fill_value, dtype=x.dtype).reshape(-1)
    if fill.numel() != 1:
        raise MexError(f"{name}:error: fill_value should contain exactly one element")
    nz, ny, nx = (int(d) for d in gridpos.shape[:3])
    if nz < 2 or ny < 2 or nx < 2:
        raise MexError(f"{name}:error: grid size must be at least (2, 2, 2), got "
                       f"({nx}, {ny}, {nz})")
    nv = x.dim()
    if nv < 3:
        raise MexError(f"{name}:error: x must have at least 3 dimensions")
    if tuple(x.shape[nv - 3:]) != (nz, ny, nx):
        raise MexError(f"{name}:error: x spatial dims must match grid shape ({nx}, {ny}, {nz})")
    nq = querypos.dim()
    if nq < 2 or querypos.shape[nq - 1] != 3:
        raise MexError(f"{name}:error: querypos must have shape (3, ...)")
    modes = {"nearest": 0, "linear": 1}
    if mode not in modes:
        raise MexError(f"{name}:error: unknown interpolation mode: {mode}")
    num_q = 1
    for d in querypos.shape[:-1]:
        num_q *= int(d)
    out = torch.empty(tuple(x.shape[:nv - 3]) + tuple(querypos.shape[:-1]), dtype=x.dtype,
                      device=x.device)
    xd = TensorDesc(tuple(x.shape), tuple(x.stride()), element_code(x))
    plan = ctypes.c_void_p()
    L = lib()
    with torch.cuda.device(x.device):
        h = handle_for(x)
        check(L.ffdas_create_interpolation_plan(h.h, ctypes.byref(plan), nx, ny, nz,
                                                gridpos.data_ptr(), modes[mode]),
              "ffdas_create_interpolation_plan")
        err = L.ffdas_interpolation(h.h, plan, num_q, querypos.data_ptr(), xd.ptr, x.data_ptr(),
                                    out.data_ptr(), fill.data_ptr())
        L.ffdas_destroy_interpolation_plan(h.h, plan)
    if err:
        raise MexError(f"{name}:error: ffdas_interpolation returned error {err}")
    return matlab_squeeze(out)


def mex_greens(srcpos, wavenums, x, dstpos):
    """greens.m + ffdas_greens.cpp:13-72: x (frequencies, sources[, batch]) in MATLAB, i.e.
    the row-major (batch, sources, frequencies) view, a squeezed batch re-inserted; the
    output (frequencies, dst..., batch), squeezed as MATLAB does."""
    name = "ffdas_greens"
    if x.dim() == 2:
        x = x.reshape((1,) + tuple(x.shape))
    elif x.dim() != 3:
        raise MexError(f"{name}:error: input must have 2 or 3 dimensions")
    B, ns, nf = (int(d) for d in x.shape)
    dims = tuple(int(d) for d in dstpos.shape[:-1])
    ndst = 1
    for d in dims:
        ndst *= d
    out = torch.empty((B, ndst, nf), dtype=x.dtype, device=x.device)
    xd = TensorDesc((B, ns, nf), tuple(x.stride()), element_code(x))
    od = TensorDesc((B, ndst, nf), tuple(out.stride()), element_code(out))
    with torch.cuda.device(x.device):
        h = handle_for(x)
        check(lib().ffdas_greens(h.h, srcpos.data_ptr(), wavenums.data_ptr(), xd.ptr,
                                 x.data_ptr(), dstpos.data_ptr(), od.ptr, out.data_ptr()), name)
    return matlab_squeeze(out.reshape((B,) + dims + (nf,)))
