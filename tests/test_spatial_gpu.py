"""Transmit-geometry helpers on ROCm tensors (GPU): cdist / rect_dist / angle (reference
bindings/python/ffdas/utils/spatial.py:10-50) against closed forms, and the quickstart's
diverging-wave and plane-wave delays and angular apodization built with them on the device
and fed into ``das`` against the FP64 oracle (docs/getting-started/quickstart.md:44-81)."""
import math

import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_cdist_on_device(dev):
    import torch
    from ffdas_amd.utils import spatial
    g = torch.Generator(device="cpu").manual_seed(0)
    a = torch.randn((4, 5, 3), generator=g, dtype=torch.float64).to(dev)
    b = torch.randn((7, 3), generator=g, dtype=torch.float64).to(dev)
    d = spatial.cdist(a, b)
    assert d.device == a.device and d.shape == (4, 5, 7)
    want = torch.cdist(a.reshape(-1, 3), b).reshape(4, 5, 7)
    assert torch.allclose(d, want, rtol=1e-12, atol=1e-12)
    # unit offsets along the axes: distance 1 exactly; Pythagorean triple: 5
    o = torch.zeros((1, 3), dtype=torch.float64, device=dev)
    e = torch.tensor([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [3.0, 4.0, 0]],
                     dtype=torch.float64, device=dev)
    assert torch.equal(spatial.cdist(o, e)[0], torch.tensor([1.0, 1.0, 1.0, 5.0],
                                                            dtype=torch.float64, device=dev))


def test_rect_dist_on_device(dev):
    import torch
    from ffdas_amd.utils import spatial
    size = torch.tensor([4.0, 2.0], dtype=torch.float64, device=dev)
    pts = torch.tensor([[0.0, 0.0, 3.0], [5.0, 0.0, 0.0], [5.0, 4.0, 0.0], [1.0, 0.5, -2.0],
                        [0.3, 0.3, 2.0], [2.0, 0.0, 0.0]], dtype=torch.float64, device=dev)
    want = torch.tensor([3.0, 3.0, math.hypot(3.0, 3.0), 2.0, 2.0, 0.0], dtype=torch.float64)
    got = spatial.rect_dist(pts, size)
    assert got.device == pts.device
    assert torch.allclose(got.cpu(), want, rtol=0, atol=1e-12)


def test_angle_on_device(dev):
    import torch
    from ffdas_amd.utils import spatial
    a = torch.tensor([[1.0, 0, 0]] * 4, dtype=torch.float64, device=dev)
    b = torch.tensor([[2.0, 0, 0], [0.0, 1, 0], [-1.0, 0, 0], [1.0, 1.0, 0]],
                     dtype=torch.float64, device=dev)
    got = spatial.angle(a, b).cpu().numpy()
    assert np.allclose(got, [0.0, math.pi / 2, math.pi, math.pi / 4], atol=1e-3)
    assert np.allclose(got[[1, 3]], [math.pi / 2, math.pi / 4], atol=1e-6)


def test_diverging_wave_das_from_device_geometry(dev):
    """A small volumetric problem whose transmit delays and weights are produced on the GPU by
    the helpers: 8x8 matrix array, 4 virtual sources 10 mm behind it, 20^3 voxels, directivity
    cone of 30 degrees, Hann window of the source angle (f-number 1).  The delays equal the
    workload generator's formula, and das on them matches the FP64 oracle."""
    import torch
    import ffdas_amd
    from ffdas_amd.utils import spatial
    elem = W.matrix_array(8, 0.3e-3)
    src = np.array([[-1.5e-3, -1.5e-3, -10e-3], [1.5e-3, -1.5e-3, -10e-3],
                    [-1.5e-3, 1.5e-3, -10e-3], [1.5e-3, 1.5e-3, -10e-3]])
    lat = np.linspace(-4e-3, 4e-3, 20)
    dep = np.linspace(5e-3, 20e-3, 20)
    zz, yy, xx = np.meshgrid(dep, lat, lat, indexing="ij")
    grid = np.stack([xx, yy, zz], axis=-1)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    g_d, s_d = t(grid), t(src)
    ap = torch.tensor([8 * 0.3e-3, 8 * 0.3e-3], dtype=torch.float64, device=dev)
    ref = spatial.rect_dist(s_d, ap)                                        # (4,)
    offs = (spatial.cdist(s_d, g_d) - ref[:, None, None, None]) * W.KS     # (4, 20, 20, 20)
    assert np.allclose(offs.cpu().numpy(), W.diverging_delay(grid, src, ref.cpu().numpy()),
                       rtol=0, atol=1e-9)
    vec = g_d[None] - s_d[:, None, None, None, :]
    zax = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64, device=dev).expand_as(vec)
    th = spatial.angle(vec, zax)
    ratio = torch.tan(th)
    wts = torch.where(ratio.abs() <= 0.5, 0.5 + 0.5 * torch.cos(2 * math.pi * ratio),
                      torch.zeros_like(ratio))
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 64, 4, 1024)).astype(np.float32)
    srcdir = np.tile(np.array([0, 0, 1, math.cos(math.radians(30.0))], np.float32), (64, 1))
    args = dict(srcpos=t((elem * W.KS).astype(np.float32)),
                dstpos=t((grid * W.KS).astype(np.float32)))
    out = ffdas_amd.das(t(x), args["srcpos"], args["dstpos"], offs.float(), wts.float(),
                        srcdir=t(srcdir))
    want = oracle.das(x, (elem * W.KS).astype(np.float32), (grid * W.KS).astype(np.float32),
                      offs.float().cpu().numpy(), wts.float().cpu().numpy(), srcdir=srcdir)
    assert rel_l2(out.cpu().numpy(), want) < 1e-5


def test_plane_wave_das_from_device_geometry(dev):
    """Plane-wave delays z cos(theta) + x sin(theta) computed on the device for a linear array
    and fed into das (FP16 storage) against the FP64 oracle on the rounded data."""
    import torch
    import ffdas_amd
    elem = W.linear_array(32, 0.3e-3)
    xs = np.linspace(-5e-3, 5e-3, 24)
    zs = np.linspace(5e-3, 25e-3, 40)
    zz, xx = np.meshgrid(zs, xs, indexing="ij")
    grid = np.stack([xx, np.zeros_like(xx), zz], axis=-1)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    g_d = t(grid)
    th = torch.deg2rad(torch.linspace(-10.0, 10.0, 5, dtype=torch.float64, device=dev))
    offs = (g_d[None, ..., 2] * torch.cos(th)[:, None, None] +
            g_d[None, ..., 0] * torch.sin(th)[:, None, None]) * W.KS
    assert np.allclose(offs.cpu().numpy(), W.plane_wave_delay(grid, th.cpu().numpy()),
                       rtol=0, atol=1e-9)
    wts = torch.full(offs.shape, 0.2, dtype=torch.float32, device=dev)
    x = np.random.default_rng(4).standard_normal((3, 32, 5, 1400)).astype(np.float32)
    out = ffdas_amd.das(t(x), t((elem * W.KS).astype(np.float32)),
                        t((grid * W.KS).astype(np.float32)), offs.float(), wts, use_fp16=True)
    xq = x.astype(np.float16).astype(np.float64)
    want = oracle.das(xq, (elem * W.KS).astype(np.float32), (grid * W.KS).astype(np.float32),
                      offs.float().cpu().numpy(), wts.cpu().numpy())
    assert rel_l2(out.cpu().numpy(), want) < 2e-5
