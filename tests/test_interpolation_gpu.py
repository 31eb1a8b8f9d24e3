"""interpolate / Interpolator on the GPU against the oracle (oracle/interpolation.py): the
reference's own cases (bindings/python/tests/test_interpolation.py:15-104), every dtype with a
frame axis, points inside, outside and on a curvilinear sector grid, the preprocessed path,
the C-ABI's error codes and an affine field on a large grid (a size-independent check).

The GPU computes the bins, the barycentric coordinates and the weighted sums in FP32 (FP64
products for double data) in the oracle's operation order without contraction, so the
results are expected to agree to the last bit; the comparisons allow 1e-6 relative."""
import ctypes

import numpy as np
import pytest

import oracle.interpolation as I

pytestmark = pytest.mark.gpu


def _regular(nz, ny, nx):
    z = np.linspace(0, 1, nz, dtype="float32")
    y = np.linspace(0, 1, ny, dtype="float32")
    x = np.linspace(0, 1, nx, dtype="float32")
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([xx, yy, zz], axis=-1)


def _sector(nr, nt, np_, r0=1.0, r1=2.0):
    r = np.linspace(r0, r1, nr)
    th = np.linspace(-0.5, 0.5, nt)
    ph = np.linspace(-0.4, 0.4, np_)
    R, TH, PH = np.meshgrid(r, th, ph, indexing="ij")
    return np.stack([R * np.sin(TH) * np.cos(PH), R * np.sin(PH),
                     R * np.cos(TH) * np.cos(PH)], -1).astype(np.float32)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_reference_exact_at_grid_points(dev, mode):
    import ffdas_amd
    g = _regular(3, 4, 5)
    x = np.arange(60, dtype=np.float32).reshape(3, 4, 5)
    interp = ffdas_amd.Interpolator(_t(g, dev), mode=mode)
    r = interp(_t(x, dev), _t(g.reshape(-1, 3), dev))
    np.testing.assert_allclose(r.cpu().numpy(), x.ravel(), atol=1e-4)


def test_reference_linear_midpoint(dev):
    import ffdas_amd
    x = np.zeros((2, 2, 2), np.float32)
    x[0, 0, 1] = 1.0
    r = ffdas_amd.Interpolator(_t(_regular(2, 2, 2), dev))(
        _t(x, dev), _t(np.array([[0.5, 0, 0]], np.float32), dev))
    np.testing.assert_allclose(r.cpu().numpy(), [0.5], atol=1e-4)


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_reference_fill_outside(dev, mode):
    import ffdas_amd
    r = ffdas_amd.Interpolator(_t(_regular(2, 2, 2), dev), mode=mode)(
        _t(np.ones((2, 2, 2), np.float32), dev), _t(np.array([[5, 5, 5.0]], np.float32), dev),
        fill=-1.0)
    np.testing.assert_allclose(r.cpu().numpy(), [-1.0], atol=1e-5)


def test_reference_output_shape(dev):
    import ffdas_amd
    r = ffdas_amd.Interpolator(_t(_regular(2, 3, 4), dev))(
        _t(np.ones((2, 3, 4), np.float32), dev), _t(np.zeros((10, 3), np.float32), dev))
    assert tuple(r.shape) == (10,)


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_reference_batch_dim(dev, mode):
    import ffdas_amd
    g = _regular(3, 4, 5)
    x = np.arange(4 * 60, dtype=np.float32).reshape(4, 3, 4, 5)
    r = ffdas_amd.Interpolator(_t(g, dev), mode=mode)(_t(x, dev), _t(g.reshape(-1, 3), dev))
    np.testing.assert_allclose(r.cpu().numpy(), x.reshape(4, -1), atol=1e-4)


def test_reference_convenience_and_dims_mismatch(dev):
    import ffdas_amd
    g = _regular(3, 4, 5)
    x = np.arange(60, dtype=np.float32).reshape(3, 4, 5)
    r = ffdas_amd.interpolate(_t(g, dev), _t(x, dev), _t(g.reshape(-1, 3), dev), mode="nearest")
    np.testing.assert_allclose(r.cpu().numpy(), x.ravel(), atol=1e-5)
    interp = ffdas_amd.Interpolator(_t(_regular(2, 3, 4), dev))
    with pytest.raises(ValueError):
        interp(_t(np.ones((3, 3, 4), np.float32), dev), _t(np.zeros((1, 3), np.float32), dev))


@pytest.mark.parametrize("mode", ["nearest", "linear"])
@pytest.mark.parametrize("dtype", ["float32", "complex64", "float64", "complex128"])
def test_sector_grid_against_oracle(dev, mode, dtype):
    """A curvilinear (spherical sector) grid, 3 frames, queries inside, outside and beyond
    the bin range; the oracle's result including which points are filled."""
    import ffdas_amd
    g = _sector(7, 9, 6)
    rng = np.random.default_rng(41)
    shape = (3,) + g.shape[:3]
    x = rng.standard_normal(shape)
    if dtype.startswith("complex"):
        x = x + 1j * rng.standard_normal(shape)
    x = x.astype(dtype)
    flat = g.reshape(-1, 3)
    q = rng.uniform(flat.min(0) - 0.2, flat.max(0) + 0.2, (40, 50, 3)).astype(np.float32)
    fill = np.nan
    want = I.interpolate(g, x, q, mode=mode, fill=fill)
    got = ffdas_amd.Interpolator(_t(g, dev), mode=mode)(_t(x, dev), _t(q, dev), fill=fill)
    got = got.cpu().numpy()
    assert got.shape == want.shape == (3, 40, 50)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6, equal_nan=True)
    assert 0.2 < np.mean(~np.isnan(want)) < 0.9


def test_preprocessed_plan_is_reused(dev):
    import torch
    import ffdas_amd
    g = _sector(6, 7, 5)
    rng = np.random.default_rng(42)
    flat = g.reshape(-1, 3)
    q = _t(rng.uniform(flat.min(0), flat.max(0), (500, 3)).astype(np.float32), dev)
    x1 = _t(rng.standard_normal((2,) + g.shape[:3]).astype(np.float32), dev)
    x2 = x1 * 3.0 + 1.0
    interp = ffdas_amd.Interpolator(_t(g, dev))
    a = interp(x1, q, preprocess=True)
    b = interp(x2, q)                         # same number of points: the cached lookup
    assert torch.equal(a, ffdas_amd.interpolate(_t(g, dev), x1, q))
    assert torch.equal(b, ffdas_amd.interpolate(_t(g, dev), x2, q))
    out = torch.empty_like(b)
    assert interp(x2, q, out=out) is out and torch.equal(out, b)


def test_flat_grid_fills_everything(dev):
    """A grid with one layer has no cells: every point gets the fill value."""
    import ffdas_amd
    g = _regular(1, 4, 5)
    r = ffdas_amd.interpolate(_t(g, dev), _t(np.ones((1, 4, 5), np.float32), dev),
                              _t(g.reshape(-1, 3), dev), fill=7.0)
    assert np.all(r.cpu().numpy() == 7.0)


def test_affine_field_on_a_large_grid(dev):
    """Size-independent check at a scan-conversion size (96 x 64 x 64 sector to 64^3 points,
    4 frames): linear interpolation reproduces an affine field at every located point."""
    import torch
    import ffdas_amd
    g = _sector(96, 64, 64, r0=0.2, r1=2.0)
    coef = np.array([0.3, -1.2, 0.7], np.float32)
    base = (g @ coef + 0.1).astype(np.float32)
    x = np.stack([base * (k + 1) for k in range(4)])
    flat = g.reshape(-1, 3)
    ax = [np.linspace(flat[:, a].min(), flat[:, a].max(), 64, dtype=np.float32) for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    q = np.stack([X, Y, Z], -1).astype(np.float32)
    r = ffdas_amd.Interpolator(_t(g, dev))(_t(x, dev), _t(q, dev), fill=float("nan"))
    r = r.cpu().numpy().reshape(4, -1)
    want = q.reshape(-1, 3) @ coef + 0.1
    ok = ~np.isnan(r[0])
    assert ok.mean() > 0.2
    for k in range(4):
        assert np.abs(r[k, ok] - (k + 1) * want[ok]).max() < 1e-5 * (k + 1)


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_locate_by_bins_matches_per_point_scan(dev, monkeypatch, mode):
    """Location by bins (points ordered by bin, one wave per bin testing the bin's candidates
    once for all its points; the default) and the per-point scan (FFDAS_INTERP_LOCATE=point)
    pick the same tetrahedron, the lowest passing id of the point's bin, so the interpolated
    values agree bit for bit.  The sector grid is fine along range (about 4.7 cells per bin),
    so bins hold more candidates than one wave of 64; queries are spread inside and outside,
    on grid vertices (shared by several tetrahedra: the id order decides), and clustered in a
    few bins (many points per bin)."""
    import torch
    import ffdas_amd
    g = _sector(600, 20, 20, r0=0.3, r1=2.0)
    rng = np.random.default_rng(44)
    flat = g.reshape(-1, 3)
    lo, hi = flat.min(0), flat.max(0)
    centre = flat[rng.integers(0, flat.shape[0])]
    cluster = centre + rng.uniform(-1.0, 1.0, (5000, 3)) * (hi - lo) / 128
    q = np.concatenate([rng.uniform(lo - 0.1, hi + 0.1, (40000, 3)),
                        flat[rng.integers(0, flat.shape[0], 5000)], cluster]).astype(np.float32)
    x = _t(rng.standard_normal((2,) + g.shape[:3]).astype(np.float32), dev)
    interp = ffdas_amd.Interpolator(_t(g, dev), mode=mode)
    monkeypatch.setenv("FFDAS_INTERP_LOCATE", "point")
    a = interp(x, _t(q, dev), fill=float("nan"))
    monkeypatch.delenv("FFDAS_INTERP_LOCATE")
    b = interp(x, _t(q, dev), fill=float("nan"))
    assert torch.equal(torch.isnan(a), torch.isnan(b))
    assert torch.equal(torch.nan_to_num(a), torch.nan_to_num(b))
    assert 0.2 < float((~torch.isnan(b[0])).float().mean()) < 0.95


def test_abi_error_codes(dev):
    import torch
    import ffdas_amd
    from ffdas_amd._core import handle_for
    from ffdas_amd._lib import TensorDesc, lib
    L = lib()
    g = _t(_regular(2, 3, 4), dev)
    h = handle_for(g)
    plan = ctypes.c_void_p()
    INVALID_ARGUMENT, INVALID_DIMS, UNSUPPORTED_TYPE, INVALID_HANDLE = 1, 2, 3, 5
    assert L.ffdas_create_interpolation_plan(None, ctypes.byref(plan), 4, 3, 2, g.data_ptr(), 1) == INVALID_HANDLE
    assert L.ffdas_create_interpolation_plan(h.h, ctypes.byref(plan), 0, 3, 2, g.data_ptr(), 1) == INVALID_DIMS
    assert L.ffdas_create_interpolation_plan(h.h, ctypes.byref(plan), 4, 3, 2, g.data_ptr(), 7) == INVALID_ARGUMENT
    assert L.ffdas_create_interpolation_plan(h.h, ctypes.byref(plan), 4, 3, 2, g.data_ptr(), 1) == 0
    try:
        x = torch.ones((2, 3, 4), device=dev)
        out = torch.empty(5, device=dev)
        fill = torch.zeros(1)
        q = torch.zeros((5, 3), device=dev)
        xd = TensorDesc((2, 3, 4), (12, 4, 1), 0)
        assert L.ffdas_interpolation(h.h, plan, 5, None, xd.ptr, x.data_ptr(), out.data_ptr(),
                                     fill.data_ptr()) == INVALID_ARGUMENT
        assert L.ffdas_interpolation(h.h, plan, 0, q.data_ptr(), xd.ptr, x.data_ptr(),
                                     out.data_ptr(), fill.data_ptr()) == INVALID_DIMS
        bad = TensorDesc((3, 3, 4), (12, 4, 1), 0)
        assert L.ffdas_interpolation(h.h, plan, 5, q.data_ptr(), bad.ptr, x.data_ptr(),
                                     out.data_ptr(), fill.data_ptr()) == INVALID_DIMS
        xh = x.half()
        hd = TensorDesc((2, 3, 4), (12, 4, 1), 2)
        assert L.ffdas_interpolation(h.h, plan, 5, q.data_ptr(), hd.ptr, xh.data_ptr(),
                                     out.data_ptr(), fill.data_ptr()) == UNSUPPORTED_TYPE
        assert L.ffdas_interpolation(h.h, plan, 5, q.data_ptr(), xd.ptr, x.data_ptr(),
                                     out.data_ptr(), fill.data_ptr()) == 0
        torch.cuda.synchronize()
    finally:
        assert L.ffdas_destroy_interpolation_plan(h.h, plan) == 0
