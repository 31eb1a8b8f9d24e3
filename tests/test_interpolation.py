"""The interpolation oracle (oracle/interpolation.py) pinned on the CPU, and the Python
wrapper's argument errors (no GPU needed).

Pins: the reference's own cases (bindings/python/tests/test_interpolation.py:15-104: exact at
grid points, the midpoint, fill outside, batch, output shape); linear mode reproduces affine
fields exactly up to FP32 rounding on a curvilinear grid (a closed form that holds whichever
tetrahedron is chosen); nearest mode returns the nearest vertex of a regular grid away from
ties; the binned search equals a brute-force search over all tetrahedra; and the six
tetrahedra of a cell tile it (their volumes sum to the cell's)."""
import numpy as np
import pytest

import oracle.interpolation as I


def _regular(nz, ny, nx):
    z = np.linspace(0, 1, nz, dtype="float32")
    y = np.linspace(0, 1, ny, dtype="float32")
    x = np.linspace(0, 1, nx, dtype="float32")
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([xx, yy, zz], axis=-1)


def _sector(nr=6, nt=7, np_=5):
    r = np.linspace(1, 2, nr)
    th = np.linspace(-0.5, 0.5, nt)
    ph = np.linspace(-0.4, 0.4, np_)
    R, TH, PH = np.meshgrid(r, th, ph, indexing="ij")
    return np.stack([R * np.sin(TH) * np.cos(PH), R * np.sin(PH),
                     R * np.cos(TH) * np.cos(PH)], -1).astype(np.float32)


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_reference_exact_at_grid_points(mode):
    g = _regular(3, 4, 5)
    x = np.arange(60, dtype=np.float32).reshape(3, 4, 5)
    np.testing.assert_allclose(I.interpolate(g, x, g.reshape(-1, 3), mode=mode), x.ravel(),
                               atol=1e-4)


def test_reference_linear_midpoint():
    g = _regular(2, 2, 2)
    x = np.zeros((2, 2, 2), np.float32)
    x[0, 0, 1] = 1.0
    np.testing.assert_allclose(I.interpolate(g, x, np.array([[0.5, 0, 0]], np.float32)), [0.5],
                               atol=1e-4)


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_reference_fill_outside(mode):
    g = _regular(2, 2, 2)
    r = I.interpolate(g, np.ones((2, 2, 2), np.float32), np.array([[5, 5, 5.0]], np.float32),
                      mode=mode, fill=-1.0)
    np.testing.assert_allclose(r, [-1.0])


@pytest.mark.parametrize("mode", ["nearest", "linear"])
def test_reference_batch(mode):
    g = _regular(3, 4, 5)
    x = np.arange(4 * 60, dtype=np.float32).reshape(4, 3, 4, 5)
    r = I.interpolate(g, x, g.reshape(-1, 3), mode=mode)
    assert r.shape == (4, 60)
    np.testing.assert_allclose(r, x.reshape(4, -1), atol=1e-4)


def test_linear_reproduces_affine_fields():
    g = _sector()
    vals = (0.3 * g[..., 0] - 1.2 * g[..., 1] + 0.7 * g[..., 2] + 0.1).astype(np.float32)
    rng = np.random.default_rng(0)
    flat = g.reshape(-1, 3)
    q = rng.uniform(flat.min(0), flat.max(0), (400, 3)).astype(np.float32)
    out = I.interpolate(g, vals, q, mode="linear", fill=np.nan)
    ok = ~np.isnan(out)
    assert ok.sum() > 100                        # the sector fills about half its box
    want = 0.3 * q[:, 0] - 1.2 * q[:, 1] + 0.7 * q[:, 2] + 0.1
    assert np.abs(out[ok] - want[ok]).max() < 2e-6


def test_nearest_returns_the_vertex_a_point_sits_next_to():
    """A point within 5 % of a cell from a grid vertex has that vertex's barycentric weight
    above 0.85 in any tetrahedron holding it, so nearest mode must return it."""
    g = _sector()
    x = np.arange(g.shape[0] * g.shape[1] * g.shape[2], dtype=np.float32).reshape(g.shape[:3])
    rng = np.random.default_rng(1)
    flat = g.reshape(-1, 3)
    pick = rng.integers(0, flat.shape[0], 200)
    cell = 0.05 * 0.15                                # 5 % of the smallest cell edge
    q = (flat[pick] + rng.uniform(-cell, cell, (200, 3)) / np.sqrt(3)).astype(np.float32)
    got = I.interpolate(g, x, q, mode="nearest", fill=-1.0)
    located = got >= 0
    assert located.mean() > 0.3                       # points off the boundary faces
    np.testing.assert_array_equal(got[located], pick[located].astype(np.float32))


def test_binned_search_equals_brute_force():
    g = _sector(5, 6, 4)
    p = I.Plan(g)
    rng = np.random.default_rng(2)
    flat = g.reshape(-1, 3)
    q = rng.uniform(flat.min(0) - 0.1, flat.max(0) + 0.1, (300, 3)).astype(np.float32)
    for i in range(q.shape[0]):
        assert p.locate(q[i])[0] == p.locate(q[i], brute=True)[0]


def test_cell_split_partitions_the_cell():
    """The reference's split of a cell is four corner tetrahedra of volume 1/6, a central one
    of volume 1/3 and a degenerate sixth (the top face) that never holds a point: the absolute
    volumes sum to the cell and every point of the cell lies in exactly one tetrahedron."""
    cube = _regular(2, 2, 2)
    pc = I.Plan(cube)
    vc = pc.v[pc.vid]
    vol = np.abs(np.einsum("ij,ij->i", vc[:, 1] - vc[:, 0],
                           np.cross(vc[:, 2] - vc[:, 0], vc[:, 3] - vc[:, 0]))) / 6
    np.testing.assert_allclose(vol, [1 / 6, 1 / 6, 1 / 3, 1 / 6, 1 / 6, 0], atol=1e-7)
    rng = np.random.default_rng(3)
    for q in rng.uniform(0, 1, (500, 3)).astype(np.float32):
        _, ok = pc.barycentric(q, np.arange(6))
        assert ok.sum() == 1


def test_wrapper_argument_errors():
    torch = pytest.importorskip("torch")
    import ffdas_amd
    with pytest.raises(ValueError):
        ffdas_amd.Interpolator(torch.zeros((4, 3)))
    with pytest.raises(ValueError):
        ffdas_amd.Interpolator(torch.zeros((2, 3, 4, 2)))
    with pytest.raises(ValueError):
        ffdas_amd.Interpolator(torch.zeros((2, 3, 4, 3)), mode="cubic")
