"""The interpolate and greens MEX gateways' conventions (tests/mex_gateway.py, restating
ffdas_interpolate.cpp and ffdas_greens.cpp) over the C-ABI: results equal the Python API's
on the same row-major views, MATLAB's squeezed singleton dimensions are handled, and the
gateways' argument errors are raised."""
import numpy as np
import pytest

from tests.mex_gateway import MexError, mex_greens, mex_interpolate

pytestmark = pytest.mark.gpu


def _grid(nz, ny, nx):
    z, y, x = (np.linspace(0, 1, n, dtype="float32") for n in (nz, ny, nx))
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([xx, yy, zz], -1)


def test_interpolate_gateway(dev):
    import torch
    import ffdas_amd
    g = torch.from_numpy(_grid(3, 4, 5)).to(dev)
    rng = np.random.default_rng(50)
    x = torch.from_numpy(rng.standard_normal((2, 3, 4, 5)).astype("float32")).to(dev)
    q = torch.from_numpy(rng.uniform(-0.1, 1.1, (7, 6, 3)).astype("float32")).to(dev)
    for mode in ("nearest", "linear"):
        got = mex_interpolate(g, x, q, mode, -2.0)
        assert torch.equal(got, ffdas_amd.interpolate(g, x, q, mode=mode, fill=-2.0))
    one = mex_interpolate(g, x[0], q, "linear", 0.0)          # MATLAB x (5, 4, 3)
    assert tuple(one.shape) == (7, 6)
    assert torch.equal(one, ffdas_amd.interpolate(g, x[0], q, fill=0.0))
    with pytest.raises(MexError):
        mex_interpolate(g[..., :2], x, q, "linear", 0.0)
    with pytest.raises(MexError):
        mex_interpolate(g, x, q, "cubic", 0.0)
    with pytest.raises(MexError):
        mex_interpolate(g, x, q, "linear", [0.0, 1.0])
    with pytest.raises(MexError):
        mex_interpolate(torch.from_numpy(_grid(1, 4, 5)).to(dev), x[:, :1], q, "linear", 0.0)
    with pytest.raises(MexError):
        mex_interpolate(g, x[:, :2], q, "linear", 0.0)


def test_greens_gateway(dev):
    import torch
    import ffdas_amd
    rng = np.random.default_rng(51)
    src = torch.from_numpy(rng.standard_normal((6, 3)).astype("float32")).to(dev)
    dst = torch.from_numpy((rng.standard_normal((4, 5, 3)) + 3).astype("float32")).to(dev)
    k = torch.from_numpy(rng.uniform(1, 10, 9).astype("float32")).to(dev)
    x = torch.complex(torch.randn((3, 6, 9)), torch.randn((3, 6, 9))).to(dev)
    got = mex_greens(src, k, x, dst)
    assert tuple(got.shape) == (3, 4, 5, 9)
    assert torch.equal(got, ffdas_amd.greens(src, k, x, dst))
    one = mex_greens(src, k, x[0], dst)                       # MATLAB x (9, 6): no batch
    assert tuple(one.shape) == (4, 5, 9)
    assert torch.equal(one, ffdas_amd.greens(src, k, x[0], dst))
    with pytest.raises(MexError):
        mex_greens(src, k, x[None], dst)
