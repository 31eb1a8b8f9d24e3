"""Transmit-geometry helpers (CPU): cdist / rect_dist / angle against closed forms and the
workload generators' own delay formulas."""
import math

import numpy as np
import torch

from ffdas_amd import workloads as W
from ffdas_amd.utils import spatial


def test_cdist_shapes_and_values():
    a = np.random.default_rng(0).standard_normal((4, 5, 3))
    b = np.random.default_rng(1).standard_normal((7, 3))
    d = spatial.cdist(a, b)
    assert d.shape == (4, 5, 7)
    assert np.allclose(d[2, 3, 6], np.linalg.norm(a[2, 3] - b[6]))
    dt = spatial.cdist(torch.from_numpy(a), torch.from_numpy(b))
    assert isinstance(dt, torch.Tensor) and np.allclose(dt.numpy(), d)


def test_rect_dist_closed_forms():
    size = np.array([4.0, 2.0])
    pts = np.array([[0.0, 0.0, 3.0],      # above the rectangle: depth only
                    [5.0, 0.0, 0.0],      # beside it in x: 5 - 2
                    [5.0, 4.0, 0.0],      # corner: (3, 3)
                    [1.0, 0.5, -2.0]])    # over it, negative z
    want = np.array([3.0, 3.0, math.hypot(3.0, 3.0), 2.0])
    assert np.allclose(spatial.rect_dist(pts, size), want)
    got = spatial.rect_dist(torch.from_numpy(pts), torch.from_numpy(size))
    assert np.allclose(got.numpy(), want)


def test_rect_dist_edge_inside_and_batch_shape():
    """The reference's own cases (bindings/python/tests/test_spatial.py:47-69)."""
    assert np.allclose(spatial.rect_dist(np.array([[0.3, 0.3, 2.0]]), np.array([4.0, 4.0])), [2.0])
    assert np.allclose(spatial.rect_dist(np.array([[1.0, 0.0, 0.0]]), np.array([2.0, 2.0])), [0.0],
                       atol=1e-12)
    assert spatial.rect_dist(np.zeros((3, 4, 3)), np.array([1.0, 1.0])).shape == (3, 4)
    assert spatial.cdist(np.zeros((2, 5, 3)), np.zeros((7, 3))).shape == (2, 5, 7)


def test_angle_known_values():
    """Known angles (the reference's cases, bindings/python/tests/test_spatial.py:75-90).  The
    eps in the denominator keeps the cosine of antiparallel unit vectors at -1/(1 + 1e-7), which
    acos turns into pi - 4.5e-4: hence 1e-3 there, as in the reference's test; the other angles
    are off by at most ~1e-7."""
    a = np.array([[1.0, 0, 0], [1.0, 0, 0], [1.0, 0, 0], [1.0, 0, 0]])
    b = np.array([[2.0, 0, 0], [0.0, 1, 0], [-1.0, 0, 0], [1.0, 1.0, 0]])
    want = [0.0, math.pi / 2, math.pi, math.pi / 4]
    for got in (spatial.angle(a, b), spatial.angle(torch.from_numpy(a), torch.from_numpy(b)).numpy()):
        assert np.allclose(got, want, atol=1e-3)
        assert np.allclose(got[[1, 3]], [math.pi / 2, math.pi / 4], atol=1e-6)
    r = spatial.angle(np.ones((5, 3)), np.ones((5, 3)))
    assert r.shape == (5,) and np.allclose(r, 0.0, atol=1e-3)


def test_diverging_wave_offsets_from_the_helpers():
    """The quickstart's diverging-wave delay, (|v - y| - rect_dist(v)) * fs/c, built with the
    helpers equals the workload generator's own formula."""
    rng = np.random.default_rng(2)
    grid = rng.uniform(-0.01, 0.01, (5, 6, 3)); grid[..., 2] = np.abs(grid[..., 2]) + 0.005
    src = np.array([[0.001, -0.002, -0.01], [0.0, 0.0, -0.01]])
    ref = spatial.rect_dist(src, np.array([32 * 0.3e-3, 32 * 0.3e-3]))
    offs = (spatial.cdist(src, grid) - ref[:, None, None]) * W.KS
    assert np.allclose(offs, W.diverging_delay(grid, src, ref))
