"""truncate_rank on the GPU against the FP64 oracle (oracle/rank.py): the reference's own
cases (bindings/python/tests/test_truncate_rank.py:37-94), every dtype, both orientations
(m <= n: Gram of the rows; m > n: Gram of the columns), strided and in-place calls, more
than 128 eigenvector columns (several projection passes), and a 64 x 64^3 complex volume
ensemble built from known singular factors, so its exact truncation is known without an SVD.

Tolerances: rel-L2 against the FP64 result, 1e-4 for FP32/complex64 (the Gram matrix
squares the condition number; eigenvector errors scale with eps * s_max^2 / gap, ~1e-6 for
these Gaussian matrices) and 1e-9 for FP64/complex128."""
import numpy as np
import pytest

import oracle.rank as R
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"float32": 1e-4, "complex64": 1e-4, "float64": 1e-9, "complex128": 1e-9}


def _rand(shape, dtype, seed):
    import torch
    g = torch.Generator().manual_seed(seed)
    if dtype.is_complex:
        real = torch.float32 if dtype == torch.complex64 else torch.float64
        return torch.complex(torch.randn(shape, generator=g, dtype=real),
                             torch.randn(shape, generator=g, dtype=real)).to(dtype)
    return torch.randn(shape, generator=g, dtype=dtype)


def _check(x, start, stop, tol, **kw):
    import ffdas_amd
    got = ffdas_amd.truncate_rank(x, start, stop, **kw)
    want = R.svd_filter(x.cpu().numpy(), start, stop)
    err = rel_l2(got.cpu().numpy(), want)
    assert err < tol, err
    return got


def test_reference_keep_all(dev):
    import torch
    x = _rand((8, 32), torch.float32, 1).to(dev)
    import ffdas_amd
    got = ffdas_amd.truncate_rank(x, 0, 8)
    assert torch.equal(got, x)


@pytest.mark.parametrize("start, stop", [(0, 4), (1, 8), (2, 6)])
def test_reference_truncated(dev, start, stop):
    import torch
    _check(_rand((16, 64), torch.float32, 2).to(dev), start, stop, 1e-4)


def test_reference_output_shape_3d(dev):
    import torch
    import ffdas_amd
    x = _rand((8, 4, 16), torch.float32, 3).to(dev)
    assert ffdas_amd.truncate_rank(x, 0, 4).shape == x.shape
    _check(x, 0, 4, 1e-4)


@pytest.mark.parametrize("shape", [(16, 64), (64, 16), (32, 32)])
def test_reference_shape_variants(dev, shape):
    import torch
    k = min(shape)
    _check(_rand(shape, torch.float32, 4).to(dev), 1, k - 1, 1e-4)


def test_reference_more_rows_than_cols(dev):
    import torch
    _check(_rand((64, 8), torch.float32, 5).to(dev), 0, 8, 1e-6)


def test_reference_preallocated(dev):
    import torch
    import ffdas_amd
    x = _rand((4, 16), torch.float32, 6).to(dev)
    out = torch.zeros_like(x)
    assert ffdas_amd.truncate_rank(x, 0, out=out) is out


@pytest.mark.parametrize("dtype", ["float32", "complex64", "float64", "complex128"])
@pytest.mark.parametrize("shape, start, stop", [
    ((24, 300), 1, None),       # drop the largest component (subtract mode)
    ((24, 300), 3, 7),          # a band (keep mode)
    ((300, 20), 2, 17),         # m > n: Gram of the columns
    ((5, 7, 9), 0, 2),          # 3-D, p = 5
    ((70, 1000), 10, 60),       # p = 70: several Gram blocks
])
def test_dtypes_and_orientations(dev, dtype, shape, start, stop):
    import torch
    x = _rand(shape, getattr(torch, dtype), 7).to(dev)
    _check(x, start, stop, TOL[dtype])


def test_more_than_128_columns_and_in_place(dev):
    """r = min(kept, dropped) = 150 > 128: two projection passes; in place, the second pass
    reads a private copy of x."""
    import torch
    import ffdas_amd
    x = _rand((300, 2000), torch.float32, 8).to(dev)
    want = R.svd_filter(x.cpu().numpy(), 100, 250)
    got = ffdas_amd.truncate_rank(x, 100, 250)
    assert rel_l2(got.cpu().numpy(), want) < 1e-4
    y = x.clone()
    ffdas_amd.truncate_rank(y, 100, 250, out=y)
    assert torch.equal(y, got)


def test_in_place_single_pass(dev):
    import torch
    import ffdas_amd
    x = _rand((32, 500), torch.complex64, 9).to(dev)
    got = ffdas_amd.truncate_rank(x, 2)
    ffdas_amd.truncate_rank(x, 2, out=x)
    assert torch.equal(x, got)


def test_strided_input(dev):
    """A transposed view: element strides (1, m) -- no 16-byte loads on either side."""
    import torch
    base = _rand((400, 24), torch.complex64, 10).to(dev)
    x = base.t()                                   # (24, 400), strides (1, 24)
    _check(x, 1, 20, 1e-4)
    base = _rand((30, 410), torch.float32, 11).to(dev)
    _check(base.t(), 3, None, 1e-4)                # (410, 30): m > n, strided


def test_idempotent_and_deterministic(dev):
    import torch
    import ffdas_amd
    x = _rand((40, 3000), torch.float32, 12).to(dev)
    a = ffdas_amd.truncate_rank(x, 4, 30)
    b = ffdas_amd.truncate_rank(x, 4, 30)
    assert torch.equal(a, b)
    c = ffdas_amd.truncate_rank(a, 0, 26)          # a has rank 26: keeping it all is a
    assert float((c - a).norm() / a.norm()) < 1e-4


def test_complement_adds_up(dev):
    import torch
    import ffdas_amd
    x = _rand((50, 800), torch.complex128, 13).to(dev)
    a = ffdas_amd.truncate_rank(x, 0, 7)
    b = ffdas_amd.truncate_rank(x, 7)
    assert float((a + b - x).norm() / x.norm()) < 1e-10


def test_clutter_volume_known_factors(dev):
    """64 frames of a 64^3 complex volume (the clutter-filter example's size), x = U S V^H
    with known factors and a tissue-like spectrum (five strong components); start = 5 removes
    them.  The expected result is U[:, 5:] S[5:] V[:, 5:]^H, no SVD needed."""
    import torch
    import ffdas_amd
    m, n = 64, 64 ** 3
    g = torch.Generator().manual_seed(14)
    u, _ = torch.linalg.qr(torch.randn(m, m, dtype=torch.complex128, generator=g))
    v, _ = torch.linalg.qr(torch.randn(n, m, dtype=torch.complex128, generator=g))
    s = torch.cat([torch.tensor([4000.0, 2500.0, 1600.0, 900.0, 500.0]),
                   torch.linspace(40.0, 1.0, m - 5)]).to(torch.complex128)
    x = ((u * s) @ v.conj().T).to(dev)
    want = ((u[:, 5:] * s[5:]) @ v[:, 5:].conj().T).to(dev)
    got = ffdas_amd.truncate_rank(x.to(torch.complex64).reshape(m, 64, 64, 64), 5)
    err = float((got.reshape(m, n).to(torch.complex128) - want).norm() / want.norm())
    assert err < 1e-3, err           # FP32 Gram: 4000^2 / 40^2 condition of the kept part
    got64 = ffdas_amd.truncate_rank(x, 5)
    assert float((got64 - want).norm() / want.norm()) < 1e-9


def test_matlab_gateway(dev):
    """truncate_rank.m / ffdas_truncate_rank.cpp: 1-based start, inclusive stop, last MATLAB
    dimension kept; the same numbers as the Python call on the row-major view."""
    import torch
    import ffdas_amd
    from tests.mex_gateway import MexError, mex_truncate_rank
    x = _rand((20, 6, 7), torch.complex64, 15).to(dev)     # MATLAB (7, 6, 20)
    got = mex_truncate_rank(x, 2, 9)
    assert torch.equal(got, ffdas_amd.truncate_rank(x, 1, 9))
    assert torch.equal(mex_truncate_rank(x, 3), ffdas_amd.truncate_rank(x, 2))
    with pytest.raises(MexError):
        mex_truncate_rank(x, 0)
    with pytest.raises(MexError):
        mex_truncate_rank(x, 5, 4)


@pytest.mark.parametrize("solver", ["jacobi", "rocsolver", "tridiag"])
@pytest.mark.parametrize("dtype", ["float32", "complex64", "float64", "complex128"])
@pytest.mark.parametrize("shape, start, stop", [((33, 500), 2, 20), ((64, 2000), 5, None),
                                                 ((2, 50), 1, None)])
def test_both_eigensolvers(dev, monkeypatch, solver, dtype, shape, start, stop):
    """The one-workgroup Jacobi kernel (odd p padded with an inert index), the tridiagonal path
    (FP32; FP64 falls back to rocSOLVER) and rocSOLVER give the oracle's result."""
    import torch
    monkeypatch.setenv("FFDAS_RANK_EIG", solver)
    x = _rand(shape, getattr(torch, dtype), 16).to(dev)
    _check(x, start, stop, TOL[dtype])


def test_default_eigensolvers(dev, monkeypatch):
    """FP32 at p <= 64 takes the tridiagonal path, FP64 the Jacobi kernel."""
    import torch
    import ffdas_amd
    monkeypatch.delenv("FFDAS_RANK_EIG", raising=False)
    x = _rand((64, 4096), torch.complex64, 17).to(dev)
    _, calls = _profiled_count("rank_tridiag", lambda: ffdas_amd.truncate_rank(x, 3))
    assert calls == 1
    xd = _rand((64, 4096), torch.complex128, 17).to(dev)
    _, calls = _profiled_count("rank_jacobi", lambda: ffdas_amd.truncate_rank(xd, 3))
    assert calls == 1


@pytest.mark.parametrize("gram", ["lds", "wave"])
@pytest.mark.parametrize("dtype", ["float32", "complex64"])
@pytest.mark.parametrize("shape, start", [((40, 70001), 2), ((200, 66000), 3),
                                          ((17, 66001), 1)])
def test_both_gram_kernels(dev, monkeypatch, gram, dtype, shape, start):
    """The LDS-staged Gram kernel (FP32 default for contiguous rows, p <= 256 / 208 complex,
    at least 256 k per CU) and the per-wave block kernel, with k extents that are not a
    multiple of the 32-k stage."""
    import torch
    if gram == "wave":
        monkeypatch.setenv("FFDAS_RANK_GRAM", "wave")
    else:
        monkeypatch.delenv("FFDAS_RANK_GRAM", raising=False)
    x = _rand(shape, getattr(torch, dtype), 18).to(dev)
    _check(x, start, None, TOL[dtype])


def test_abi_error_codes(dev):
    """ffdas_truncate_rank's argument checks through the C-ABI: range, dtype, rank, shape."""
    import torch
    from ffdas_amd._core import element_code, handle_for
    from ffdas_amd._lib import TensorDesc, lib
    INVALID_ARGUMENT, INVALID_DIMS, UNSUPPORTED_TYPE = 1, 2, 3     # include/ffdas.h
    L = lib()
    x = torch.randn(8, 32, device=dev)
    out = torch.empty_like(x)
    h = handle_for(x)

    def call(xt, ot, start, stop, xdims=None, odims=None):
        xd = TensorDesc(xdims or tuple(xt.shape), None, element_code(xt))
        od = TensorDesc(odims or tuple(ot.shape), None, element_code(ot))
        return L.ffdas_truncate_rank(h.h, xd.ptr, xt.data_ptr(), start, stop, od.ptr,
                                     ot.data_ptr())

    assert call(x, out, 0, 8) == 0
    for start, stop in [(-1, 8), (8, 8), (3, 3), (2, 9)]:
        assert call(x, out, start, stop) == INVALID_ARGUMENT, (start, stop)
    xh = x.half()
    assert call(xh, torch.empty_like(xh), 0, 4) == UNSUPPORTED_TYPE
    assert call(x, out.double(), 0, 4) == UNSUPPORTED_TYPE
    assert call(x, out, 0, 4, xdims=(8, 4, 8), odims=(8, 4, 8)) == INVALID_DIMS
    assert call(x, torch.empty(8, 16, device=dev), 0, 4) == INVALID_DIMS


def test_out_aliasing_x_in_another_layout(dev):
    """out is x's own memory seen transposed: the call works from a private copy of x."""
    import torch
    from ffdas_amd._core import element_code, handle_for
    from ffdas_amd._lib import TensorDesc, lib
    import ffdas_amd
    L = lib()
    for start, stop in [(3, 40), (0, 48)]:          # subtract mode, and keep-all (a copy)
        base = _rand((48, 48), torch.float32, 19).to(dev)
        want = ffdas_amd.truncate_rank(base.clone(), start, stop)
        x = base
        out_view = base.t()                           # same memory, strides (1, 48)
        xd = TensorDesc((48, 48), (48, 1), element_code(x))
        od = TensorDesc((48, 48), (1, 48), element_code(x))
        h = handle_for(x)
        assert L.ffdas_truncate_rank(h.h, xd.ptr, x.data_ptr(), start, stop, od.ptr,
                                     out_view.data_ptr()) == 0
        torch.cuda.synchronize()
        assert float((out_view - want).norm() / want.norm()) < 1e-5, (start, stop)


@pytest.mark.parametrize("dtype", ["float32", "complex64"])
def test_power_of_two_scaling_is_exact(dev, monkeypatch, dtype):
    """Scaling x by 2^k scales every product and sum of the Gram matrix, the Jacobi
    eigensolver's input and the projection exactly (no overflow or denormals at 2^+-30), so
    the result scales exactly: truncate_rank(2^k x) == 2^k truncate_rank(x), bit for bit.
    This pins the Jacobi kernel's power-of-two prescaling of the Gram matrix.  (rocSOLVER
    rescales large or small matrices by other factors, so the property is the Jacobi
    kernel's.)"""
    import torch
    import ffdas_amd
    monkeypatch.setenv("FFDAS_RANK_EIG", "jacobi")
    x = _rand((40, 3000), getattr(torch, dtype), 21).to(dev)
    base = ffdas_amd.truncate_rank(x, 3)
    for k in (30, -30):
        got = ffdas_amd.truncate_rank(x * 2.0 ** k, 3)
        assert torch.equal(got, base * 2.0 ** k), k


@pytest.mark.parametrize("dtype", ["float32", "complex64", "float64"])
@pytest.mark.parametrize("scale", [1e12, 1e-12])
def test_extreme_magnitudes(dev, dtype, scale):
    """Magnitudes far from 1 (Gram entries near 1e24 x n or 1e-24): the oracle's result to the
    usual tolerance."""
    import torch
    x = (_rand((33, 700), getattr(torch, dtype), 22) * scale).to(dev)
    _check(x, 2, 20, TOL[dtype])


@pytest.mark.parametrize("dtype", ["float32", "complex128"])
def test_zero_input(dev, dtype):
    """All-zero data: no rotation, no component; the filtered result is zero in both the
    subtract (start 3) and keep (stop 4) forms."""
    import torch
    import ffdas_amd
    x = torch.zeros((24, 500), dtype=getattr(torch, dtype), device=dev)
    for start, stop in [(3, None), (0, 4), (1, 3)]:
        got = ffdas_amd.truncate_rank(x, start, stop)
        assert torch.equal(got, torch.zeros_like(x)), (start, stop)


@pytest.mark.parametrize("solver", ["jacobi", "rocsolver", "tridiag"])
def test_repeated_singular_values(dev, monkeypatch, solver):
    """Singular values with ties inside the removed and the kept sets (a gap only at the
    boundary): the truncation is still unique, x minus its top-3 part."""
    import torch
    import ffdas_amd
    monkeypatch.setenv("FFDAS_RANK_EIG", solver)
    m, n = 30, 4000
    g = torch.Generator().manual_seed(23)
    u, _ = torch.linalg.qr(torch.randn(m, m, dtype=torch.complex128, generator=g))
    v, _ = torch.linalg.qr(torch.randn(n, m, dtype=torch.complex128, generator=g))
    s = torch.tensor([50.0] * 3 + [5.0] * 10 + [1.0] * (m - 13), dtype=torch.complex128)
    x = ((u * s) @ v.conj().T).to(dev)
    want = ((u[:, 3:] * s[3:]) @ v[:, 3:].conj().T).to(dev)
    got = ffdas_amd.truncate_rank(x.to(torch.complex64), 3)
    assert float((got.to(torch.complex128) - want).norm() / want.norm()) < 1e-4
    got64 = ffdas_amd.truncate_rank(x, 3)
    assert float((got64 - want).norm() / want.norm()) < 1e-9
    keep = ffdas_amd.truncate_rank(x, 3, 13)      # the tied middle block alone
    want_keep = ((u[:, 3:13] * s[3:13]) @ v[:, 3:13].conj().T).to(dev)
    assert float((keep - want_keep).norm() / want_keep.norm()) < 1e-9


def _profiled_count(name, fn):
    import ctypes
    import torch
    from ffdas_amd._lib import lib
    L = lib()
    L.ffdas_profile_reset()
    L.ffdas_profile_enable(1)
    try:
        out = fn()
        torch.cuda.synchronize()
    finally:
        L.ffdas_profile_enable(0)
    t, c = ctypes.c_double(), ctypes.c_int64()
    L.ffdas_profile_query(name.encode(), ctypes.byref(t), ctypes.byref(c))
    return out, c.value


@pytest.mark.parametrize("dtype, shape, start, stop", [
    ("float32", (65, 3000), 3, None),      # the smallest p past the Jacobi kernel
    ("float32", (200, 5000), 3, None),     # the clutter filter's usual form
    ("float32", (200, 5000), 10, 150),     # keep mode: one range of 140 columns (two passes)
    ("float32", (200, 5000), 10, 180),     # subtract mode, two ranges: 20 low, 10 high columns
    ("float32", (224, 4000), 5, 218),      # the largest real p of the path, two ranges
    ("float32", (3000, 97), 2, 90),        # m > n: Gram of the columns
    ("complex64", (65, 3000), 3, None),
    ("complex64", (120, 4000), 5, None),
    ("complex64", (160, 3000), 10, 150),   # the largest complex p, subtract mode, two ranges
    ("complex64", (140, 2000), 20, 80),    # keep mode
    ("complex64", (2500, 99), 2, 90),      # m > n: the conjugated Gram of the columns
])
def test_tridiagonal_path(dev, monkeypatch, dtype, shape, start, stop):
    """FP32 with 64 < p <= 224 (real) or 160 (complex) takes the tridiagonal path by default
    (one-workgroup Householder reduction, the chosen eigenpairs by bisection and inverse
    iteration, the back-transformation of those vectors alone) and gives the oracle's
    result."""
    import torch
    import ffdas_amd
    monkeypatch.delenv("FFDAS_RANK_EIG", raising=False)
    x = _rand(shape, getattr(torch, dtype), 31).to(dev)
    got, calls = _profiled_count("rank_tridiag", lambda: ffdas_amd.truncate_rank(x, start, stop))
    assert calls == 1
    want = R.svd_filter(x.cpu().numpy(), start, stop)
    assert rel_l2(got.cpu().numpy(), want) < TOL[dtype]


def test_tridiagonal_path_limits(dev, monkeypatch):
    """p = 225 real and 161 complex (past the register tile) and FP64 keep rocSOLVER; forcing
    the path for a small p (2 and 33) gives the oracle's result too."""
    import torch
    import ffdas_amd
    monkeypatch.delenv("FFDAS_RANK_EIG", raising=False)
    x = _rand((225, 1500), torch.float32, 32).to(dev)
    got, calls = _profiled_count("rank_tridiag", lambda: ffdas_amd.truncate_rank(x, 4))
    assert calls == 0
    assert rel_l2(got.cpu().numpy(), R.svd_filter(x.cpu().numpy(), 4, None)) < TOL["float32"]
    xc = _rand((161, 1500), torch.complex64, 32).to(dev)
    got, calls = _profiled_count("rank_tridiag", lambda: ffdas_amd.truncate_rank(xc, 4))
    assert calls == 0
    assert rel_l2(got.cpu().numpy(), R.svd_filter(xc.cpu().numpy(), 4, None)) < TOL["complex64"]
    xd = _rand((100, 1500), torch.float64, 33).to(dev)
    got, calls = _profiled_count("rank_tridiag", lambda: ffdas_amd.truncate_rank(xd, 4))
    assert calls == 0
    assert rel_l2(got.cpu().numpy(), R.svd_filter(xd.cpu().numpy(), 4, None)) < TOL["float64"]
    monkeypatch.setenv("FFDAS_RANK_EIG", "tridiag")
    for shape, start, stop, dt in [((2, 50), 1, None, torch.float32),
                                   ((33, 500), 2, 20, torch.float32),
                                   ((2, 50), 1, None, torch.complex64),
                                   ((47, 900), 3, 40, torch.complex64)]:
        xs = _rand(shape, dt, 34).to(dev)
        got, calls = _profiled_count("rank_tridiag",
                                     lambda: ffdas_amd.truncate_rank(xs, start, stop))
        assert calls == 1
        assert rel_l2(got.cpu().numpy(), R.svd_filter(xs.cpu().numpy(), start, stop)) < 1e-4


@pytest.mark.parametrize("dtype", ["float32", "complex64"])
def test_tridiagonal_path_ties_scaling_zero(dev, monkeypatch, dtype):
    """The tridiagonal path with tied singular values inside the removed and kept sets (a
    gap only at the boundary; the eigenvalues of the top cluster coincide), magnitudes far
    from 1 (scaling by 2^k is exact: the kernel prescales the Gram matrix by a power of two,
    so the result scales bit for bit), and all-zero data (T = 0 splits into 1 x 1 blocks)."""
    import torch
    import ffdas_amd
    monkeypatch.delenv("FFDAS_RANK_EIG", raising=False)
    dt = getattr(torch, dtype)
    wide = torch.complex128 if dt.is_complex else torch.float64
    m, n = 100, 4000
    g = torch.Generator().manual_seed(35)

    def orth(rows, cols):
        a = torch.randn(rows, cols, dtype=torch.float64, generator=g)
        if dt.is_complex:
            a = torch.complex(a, torch.randn(rows, cols, dtype=torch.float64, generator=g))
        return torch.linalg.qr(a)[0]

    u, v = orth(m, m), orth(n, m)
    s = torch.tensor([50.0] * 3 + [5.0] * 10 + [1.0] * (m - 13), dtype=torch.float64).to(wide)
    x = (u * s) @ v.conj().T
    want = (u[:, 3:] * s[3:]) @ v[:, 3:].conj().T
    got, calls = _profiled_count("rank_tridiag",
                                 lambda: ffdas_amd.truncate_rank(x.to(dt).to(dev), 3))
    assert calls == 1
    assert float((got.cpu().to(wide) - want).norm() / want.norm()) < 1e-4
    keep = ffdas_amd.truncate_rank(x.to(dt).to(dev), 3, 13)   # the tied block
    want_keep = (u[:, 3:13] * s[3:13]) @ v[:, 3:13].conj().T
    assert float((keep.cpu().to(wide) - want_keep).norm() / want_keep.norm()) < 1e-4

    xr = _rand((120, 3000), dt, 36).to(dev)
    base = ffdas_amd.truncate_rank(xr, 3)
    for k in (30, -30):
        assert torch.equal(ffdas_amd.truncate_rank(xr * 2.0 ** k, 3), base * 2.0 ** k), k
    for scale in (1e12, 1e-12):
        xe = xr * scale
        got = ffdas_amd.truncate_rank(xe, 2, 90)
        assert rel_l2(got.cpu().numpy(), R.svd_filter(xe.cpu().numpy(), 2, 90)) < 1e-4

    z = torch.zeros((90, 500), dtype=dt, device=dev)
    for start, stop in [(3, None), (0, 4), (1, 3), (5, 80)]:
        assert torch.equal(ffdas_amd.truncate_rank(z, start, stop), torch.zeros_like(z))
