"""Host-beta accumulation through the C-ABI (GPU): ``out = beta * out + result`` with beta
read from a HOST pointer as one element of the output type (reference
src/das/das_impl.cuh:89, src/das/das_sparse_impl.cuh:85, src/das/das_device_utils.cuh:31-39).

Every store path of the library is covered: the natural-layout kernel (ALG1), the frame-inner
kernel with several waves per voxel tile (their partial sums meet in LDS before the store)
and with one wave per tile (FFDAS_CGROUPS=1), FP16 storage, complex data, FP64 data, sparse
compounding and the batched entry point (one beta per problem).  The expected value is the
FP64 oracle's ``beta * prev + das(...)`` on the same inputs; bounds as the forward tests
(FP32 1e-5, FP64 2e-6, FP16 storage 2e-5 against the oracle on the fp16-rounded data).
"""
import ctypes

import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _beta_buf(beta, out_dtype):
    import torch
    if out_dtype == torch.float64:
        return (ctypes.c_double * 1)(float(beta.real))
    if out_dtype == torch.complex128:
        return (ctypes.c_double * 2)(beta.real, beta.imag)
    if out_dtype == torch.complex64:
        return (ctypes.c_float * 2)(beta.real, beta.imag)
    return (ctypes.c_float * 1)(float(beta.real))


def _call(dev, w, x, beta, out, *, compute, alg, sparse=False):
    """ffdas_das / ffdas_das_sparse through ctypes with a host beta; out is updated in place."""
    import torch
    from ffdas_amd._core import element_code, get_library_handle
    from ffdas_amd._lib import TensorDesc, check, lib
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    srcpos, dstpos, ofs, wts = t(w.srcpos), t(w.dstpos), t(w.offsets), t(w.weights)
    srcdir = t(w.srcdir) if w.srcdir is not None else None
    xd = TensorDesc(tuple(x.shape), tuple(x.stride()), element_code(x))
    od = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
    b = _beta_buf(beta, out.dtype)
    p = lambda a: None if a is None else ctypes.c_void_p(a.data_ptr())
    h = get_library_handle(dev.index or 0)
    h.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    L = lib()
    if sparse:
        si = t(w.sparse_indices)
        rc = L.ffdas_das_sparse(h.h, p(srcpos), p(srcdir), float(w.wavenum), xd.ptr, p(x),
                                p(dstpos), p(ofs), p(wts), int(si.shape[0]), p(si),
                                ctypes.cast(b, ctypes.c_void_p), od.ptr, p(out), int(compute),
                                int(alg))
    else:
        rc = L.ffdas_das(h.h, p(srcpos), p(srcdir), float(w.wavenum), xd.ptr, p(x), p(dstpos),
                         p(ofs), p(wts), ctypes.cast(b, ctypes.c_void_p), od.ptr, p(out),
                         int(compute), int(alg))
    check(rc, "ffdas_das")
    torch.cuda.synchronize(dev)
    return out


# (x storage, compute type, algorithm, beta, complex, bound)
CASES = [
    ("f32", "FP32", "ALG1", 0.5, False, 1e-5),
    ("f32", "FP32", "DEFAULT", 0.5, False, 1e-5),
    ("f32", "FP32", "DEFAULT", -1.25, False, 1e-5),
    ("f32", "FP16", "DEFAULT", 0.5, False, 2e-5),
    ("f32", "FP16", "ALG1", 2.0, False, 2e-5),
    ("f64", "FP64", "ALG1", 0.5, False, 2e-6),
    ("c64", "FP32", "ALG1", 0.3 + 0.7j, True, 1e-5),
    ("c64", "FP32", "DEFAULT", 0.3 + 0.7j, True, 1e-5),
    ("c64", "FP16", "DEFAULT", -0.6 + 0.2j, True, 2e-5),
    ("c128", "FP64", "ALG1", 0.3 + 0.7j, True, 2e-6),
]


# the waves-per-tile override (FFDAS_CGROUPS=1) applies to the frame-inner kernels only,
# so ALG1 runs with the default groups
GROUPED = [(c, g) for c in CASES for g in (None, "1") if g is None or c[2] != "ALG1"]


@pytest.mark.parametrize("case, groups", GROUPED,
                         ids=lambda v: v if isinstance(v, str) or v is None else
                         f"{v[0]}-{v[1]}-{v[2]}-{v[3]}")
@pytest.mark.parametrize("sparse", [False, True])
def test_host_beta_accumulates(dev, monkeypatch, case, sparse, groups):
    import torch
    from ffdas_amd._lib import Algorithm, ComputeType
    kind, ct, alg, beta, cplx, bound = case
    if groups is not None:
        monkeypatch.setenv("FFDAS_CGROUPS", groups)
    w = W.small_random(seed=11, B=5, C=9, T=4, S=160, N=(7, 11), complex_iq=cplx,
                       srcdir=True, sparse=sparse, K=3 if sparse else None)
    tdt = {"f32": torch.float32, "f64": torch.float64, "c64": torch.complex64,
           "c128": torch.complex128}[kind]
    x = torch.from_numpy(w.x).to(tdt).to(dev)
    odt = {"f32": torch.float32, "f64": torch.float64, "c64": torch.complex64,
           "c128": torch.complex128}[kind]
    g = torch.Generator(device="cpu").manual_seed(3)
    prev = torch.randn((5, 7, 11), generator=g, dtype=torch.float64)
    if cplx:
        prev = prev + 1j * torch.randn((5, 7, 11), generator=g, dtype=torch.float64)
    out = prev.to(odt).to(dev)
    _call(dev, w, x, beta, out, compute=ComputeType[ct], alg=Algorithm[alg], sparse=sparse)
    xr = w.x.astype(np.complex128 if cplx else np.float64)
    if ct == "FP16":
        q = torch.from_numpy(w.x)
        q = torch.view_as_complex(torch.view_as_real(q).half().float()) if cplx else q.half()
        xr = q.numpy().astype(np.complex128 if cplx else np.float64)
    prev_np = prev.to(odt).numpy().astype(np.complex128 if cplx else np.float64)
    ref = oracle.das(xr, w.srcpos, w.dstpos, w.offsets, w.weights, srcdir=w.srcdir,
                     wavenum=w.wavenum, sparse_indices=w.sparse_indices, beta=beta,
                     out=prev_np)
    got = out.cpu().numpy()
    assert rel_l2(got, ref) < bound
    # beta really scaled the previous contents: the result differs from a plain call
    plain = oracle.das(xr, w.srcpos, w.dstpos, w.offsets, w.weights, srcdir=w.srcdir,
                       wavenum=w.wavenum, sparse_indices=w.sparse_indices)
    assert rel_l2(got, plain) > 1e-2


def test_host_beta_batched(dev):
    """ffdas_das_batched: each problem carries its own host beta (three FP16-storage problems
    of different sizes in one gather launch)."""
    import torch
    from ffdas_amd._core import element_code, get_library_handle
    from ffdas_amd._lib import Algorithm, ComputeType, DasProblem, TensorDesc, check, lib
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ws = [W.small_random(seed=20 + i, B=4, C=5 + i, T=3, S=128, N=(6 + i, 9)) for i in range(3)]
    betas = [0.5, -2.0, 0.0]
    keep, probs, outs, prevs = [], [], [], []
    g = torch.Generator(device="cpu").manual_seed(4)
    for w, beta in zip(ws, betas):
        x = t(w.x)
        prev = torch.randn((4,) + w.dstpos.shape[:-1], generator=g)
        out = prev.to(dev)
        xd = TensorDesc(tuple(x.shape), tuple(x.stride()), element_code(x))
        od = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
        b = (ctypes.c_float * 1)(beta)
        geo = [t(w.srcpos), t(w.dstpos), t(w.offsets), t(w.weights)]
        keep += [x, xd, od, b, geo]
        probs.append(DasProblem(srcpos=geo[0].data_ptr(), srcdir=None, wavenum=0.0, scale=1.0,
                                x_desc=xd.ptr, x=x.data_ptr(), dstpos=geo[1].data_ptr(),
                                offsets=geo[2].data_ptr(), weights=geo[3].data_ptr(),
                                sparse_count=0, sparse_indices=None,
                                beta=ctypes.cast(b, ctypes.c_void_p), out_desc=od.ptr,
                                out=out.data_ptr()))
        outs.append(out)
        prevs.append(prev.numpy().astype(np.float64))
    arr = (DasProblem * 3)(*probs)
    h = get_library_handle(dev.index or 0)
    h.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    check(lib().ffdas_das_batched(h.h, 3, arr, int(ComputeType.FP16), int(Algorithm.DEFAULT)),
          "ffdas_das_batched")
    torch.cuda.synchronize(dev)
    for w, beta, out, prev in zip(ws, betas, outs, prevs):
        xq = w.x.astype(np.float16).astype(np.float64)
        ref = oracle.das(xq, w.srcpos, w.dstpos, w.offsets, w.weights, beta=beta, out=prev)
        assert rel_l2(out.cpu().numpy(), ref) < 2e-5, beta
