"""MATLAB-layout callers through the C-ABI (the MEX gateways' conventions, restated in
tests/mex_gateway.py): column-major arrays seen as reversed row-major views, MATLAB's
squeezed trailing singletons, channels_trailing=false as a strided descriptor, and 1-based
sparse indices.  Outputs are compared with the FP64 oracle on the row-major problem."""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2
from tests.mex_gateway import MexError, matlab_squeeze, mex_das, mex_das_sparse

pytestmark = pytest.mark.gpu


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _matlab_x(x, channels_trailing):
    """The row-major view of the MATLAB array holding x (B, C, T, S):
    [S, T, C, B] when channels_trailing, else [S, C, T, B]; trailing singletons squeezed."""
    import torch
    t = x if channels_trailing else x.transpose(1, 2).contiguous()
    return matlab_squeeze(t)


@pytest.mark.parametrize("channels_trailing", [True, False])
@pytest.mark.parametrize("B,T", [(3, 2), (1, 3), (1, 1), (2, 1)])
def test_mex_das_layouts(dev, channels_trailing, B, T):
    w = W.small_random(seed=60 + 7 * B + T, B=B, C=5, T=T, S=90, N=(6, 7))
    x = _matlab_x(_dev(w.x, dev), channels_trailing)
    offsets = matlab_squeeze(_dev(w.offsets, dev))
    weights = matlab_squeeze(_dev(w.weights, dev))
    out = mex_das(x, _dev(w.srcpos, dev), _dev(w.dstpos, dev), offsets, weights,
                  channels_trailing=channels_trailing)
    ref = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    got = out.cpu().numpy()
    # MATLAB squeezes a leading batch of 1 from x, so the output has no batch axis then
    assert got.shape == (ref.shape if B > 1 else ref.shape[1:])
    assert rel_l2(got.reshape(ref.shape), ref) < 1e-5


def test_mex_das_fp16_strided_complex(dev):
    import torch
    w = W.small_random(seed=71, B=2, C=4, T=3, S=80, N=(5, 6), complex_iq=True)
    x = _matlab_x(_dev(w.x, dev), False)
    out = mex_das(x, _dev(w.srcpos, dev), _dev(w.dstpos, dev), _dev(w.offsets, dev),
                  _dev(w.weights, dev), wavenum=w.wavenum, use_fp16=True, channels_trailing=False)
    xq = torch.from_numpy(w.x).to(torch.complex32).to(torch.complex128).numpy()
    ref = oracle.das(xq, w.srcpos, w.dstpos, w.offsets, w.weights, wavenum=w.wavenum)
    assert out.dtype == torch.complex64
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4


@pytest.mark.parametrize("K", [3, 1])
@pytest.mark.parametrize("channels_trailing", [True, False])
def test_mex_das_sparse_one_based(dev, K, channels_trailing):
    w = W.small_random(seed=80 + K, B=2, C=5, T=4, K=K, S=90, N=(6, 7), sparse=True)
    x = _matlab_x(_dev(w.x, dev), channels_trailing)
    one_based = matlab_squeeze(_dev(w.sparse_indices + 1, dev))
    out = mex_das_sparse(x, _dev(w.srcpos, dev), _dev(w.dstpos, dev),
                         matlab_squeeze(_dev(w.offsets, dev)), matlab_squeeze(_dev(w.weights, dev)),
                         one_based, channels_trailing=channels_trailing)
    ref = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights,
                     sparse_indices=w.sparse_indices)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


def test_mex_gateway_rejects_like_the_reference(dev):
    w = W.small_random(seed=90, B=2, C=4, T=3, S=64, N=(5, 6))
    x = _dev(w.x, dev)
    sp, dp = _dev(w.srcpos, dev), _dev(w.dstpos, dev)
    of, wt = _dev(w.offsets, dev), _dev(w.weights, dev)
    with pytest.raises(MexError, match="x must have 3 or 4 dimensions"):
        mex_das(x[0, 0], sp, dp, of, wt)
    with pytest.raises(MexError, match="srcpos must have shape"):
        mex_das(x, sp[:, :2], dp, of, wt)
    with pytest.raises(MexError, match="dstpos must have shape"):
        mex_das(x, sp, dp[..., :2], of, wt)
    with pytest.raises(MexError, match="same shape"):
        mex_das(x, sp, dp, of, wt[:, :, :5])
    with pytest.raises(MexError, match="leading dimension"):
        mex_das(x, sp, dp, of[:2], wt[:2])
    with pytest.raises(MexError, match="spatial dimensions"):
        mex_das(x, sp, dp, of[:, :4], wt[:, :4])
    with pytest.raises(MexError, match="returned error"):
        mex_das(x, sp, dp, of, wt, algorithm=3)   # ALG3: INVALID_ARGUMENT from the library
