"""greens on the GPU against the FP64 oracle (oracle/greens.py): the reference's own cases
(bindings/python/tests/test_greens.py) at its tolerance (rtol 1e-3, atol 1e-4; the kernel
computes in FP32, its error is about 1e-6 of the sum's scale), all three event tilings (1
event per thread on the VALU, 4 events per thread on the VALU for 2-4 events, 16-event tiles
on the MFMA from 5 events up, each at its boundaries: 1, 2, 4, 5, 8, 9 events), partial
tiles of targets, sources, frequencies and events, FP16 input, a 2-D target grid and the
C-ABI error codes."""
import ctypes

import numpy as np
import pytest

import oracle.greens as G

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _case(rng, B, nin, nk, nout):
    src = rng.standard_normal((nin, 3)).astype("float32")
    dst = rng.standard_normal((nout, 3)).astype("float32")
    k = rng.uniform(1.0, 10.0, (nk,)).astype("float32")
    x = (rng.standard_normal((B, nin, nk)) + 1j * rng.standard_normal((B, nin, nk))).astype(
        "complex64")
    return src, k, x, dst


def test_reference_against_formula(dev):
    import ffdas_amd
    src, k, x, dst = _case(np.random.default_rng(0), 1, 4, 8, 6)
    got = ffdas_amd.greens(_t(src, dev), _t(k, dev), _t(x, dev), _t(dst, dev))
    np.testing.assert_allclose(got.cpu().numpy(), G.greens(src, k, x, dst), rtol=1e-3, atol=1e-4)


def test_reference_output_shapes(dev):
    import torch
    import ffdas_amd
    r = ffdas_amd.greens(torch.zeros((4, 3), device=dev), torch.ones(8, device=dev),
                         torch.zeros((2, 4, 8), dtype=torch.complex64, device=dev),
                         torch.zeros((6, 3), device=dev))
    assert tuple(r.shape) == (2, 6, 8)
    r = ffdas_amd.greens(torch.zeros((4, 3), device=dev), torch.ones(8, device=dev),
                         torch.zeros((1, 4, 8), dtype=torch.complex64, device=dev),
                         torch.zeros((3, 5, 3), device=dev))
    assert tuple(r.shape) == (1, 3, 5, 8)


def test_reference_zero_input(dev):
    import torch
    import ffdas_amd
    r = ffdas_amd.greens(torch.zeros((4, 3), device=dev), torch.ones(8, device=dev),
                         torch.zeros((1, 4, 8), dtype=torch.complex64, device=dev),
                         torch.ones((6, 3), device=dev))
    assert float(r.abs().max()) <= 1e-6


@pytest.mark.parametrize("B, nin, nk, nout", [(1, 37, 13, 300), (4, 64, 16, 257),
                                              (7, 20, 9, 70), (2, 1, 1, 1), (8, 37, 13, 300),
                                              (20, 66, 17, 131), (16, 3, 8, 16),
                                              (3, 29, 11, 97), (5, 33, 12, 150),
                                              (6, 40, 33, 77), (8, 1, 1, 1), (9, 21, 18, 45),
                                              (5, 3, 16, 5)])
def test_parity_with_the_oracle(dev, B, nin, nk, nout):
    import ffdas_amd
    rng = np.random.default_rng(B * 1000 + nin)
    src, k, x, dst = _case(rng, B, nin, nk, nout)
    dst = dst + 3.0                                   # keep targets off the sources
    got = ffdas_amd.greens(_t(src, dev), _t(k, dev), _t(x, dev), _t(dst, dev)).cpu().numpy()
    want = G.greens(src, k, x, dst)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() < 2e-5 * scale * max(1.0, np.sqrt(nin) / 4)


def test_two_dimensional_input_and_fp16(dev):
    import torch
    import ffdas_amd
    rng = np.random.default_rng(5)
    src, k, x, dst = _case(rng, 1, 24, 10, 40)
    dst = dst + 2.0
    got2 = ffdas_amd.greens(_t(src, dev), _t(k, dev), _t(x[0], dev), _t(dst, dev))
    assert tuple(got2.shape) == (40, 10)
    np.testing.assert_allclose(got2.cpu().numpy(), G.greens(src, k, x[0], dst), rtol=1e-3,
                               atol=1e-4)
    for xin in (x, np.concatenate([x] * 7), np.concatenate([x] * 9)):   # VALU, MFMA part/whole
        xh = _t(xin, dev).to(torch.complex32)
        goth = ffdas_amd.greens(_t(src, dev), _t(k, dev), xh, _t(dst, dev)).cpu().numpy()
        want = G.greens(src, k, xh.to(torch.complex64).cpu().numpy(), dst)
        assert np.abs(goth - want).max() < 1e-4 * np.abs(want).max()


def test_abi_error_codes(dev):
    import torch
    from ffdas_amd._core import handle_for
    from ffdas_amd._lib import TensorDesc, lib
    L = lib()
    x = torch.zeros((2, 4, 8), dtype=torch.complex64, device=dev)
    out = torch.empty((2, 6, 8), dtype=torch.complex64, device=dev)
    src, dst = torch.zeros((4, 3), device=dev), torch.ones((6, 3), device=dev)
    k = torch.ones(8, device=dev)
    h = handle_for(x)
    C32F, R32F = 4, 0
    xd = TensorDesc((2, 4, 8), (32, 8, 1), C32F)
    od = TensorDesc((2, 6, 8), (48, 8, 1), C32F)
    args = (src.data_ptr(), k.data_ptr(), xd.ptr, x.data_ptr(), dst.data_ptr(), od.ptr,
            out.data_ptr())
    assert L.ffdas_greens(None, *args) == 5
    assert L.ffdas_greens(h.h, *args) == 0
    bad = TensorDesc((2, 6, 9), (54, 9, 1), C32F)
    assert L.ffdas_greens(h.h, src.data_ptr(), k.data_ptr(), xd.ptr, x.data_ptr(),
                          dst.data_ptr(), bad.ptr, out.data_ptr()) == 2
    rd = TensorDesc((2, 4, 8), (32, 8, 1), R32F)
    assert L.ffdas_greens(h.h, src.data_ptr(), k.data_ptr(), rd.ptr, x.data_ptr(),
                          dst.data_ptr(), od.ptr, out.data_ptr()) == 3
    strided = TensorDesc((2, 4, 8), (64, 16, 2), C32F)
    assert L.ffdas_greens(h.h, src.data_ptr(), k.data_ptr(), strided.ptr, x.data_ptr(),
                          dst.data_ptr(), od.ptr, out.data_ptr()) == 6
    torch.cuda.synchronize()
