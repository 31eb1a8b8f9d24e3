"""Forward parity of the HIP library against the FP64 oracle (GPU).

Tolerances (rel-L2 against the FP64 oracle of the documented definition):
  FP32 storage      1e-5   (compensated FP32 delay arithmetic)
  FP64 storage      2e-6   (FP32 compensated geometry as in the reference, whose positions
                            are float; FP64 data and accumulation)
  FP16 storage      1e-3   against the oracle on the ORIGINAL float data; 2e-5 against the
                            oracle on the fp16-rounded data (kernel arithmetic alone)
  BF16 storage      1e-5 against the oracle on the bf16-rounded data; against the original
                   data the bf16 rounding of the data itself sets a floor of about 1e-3
                   (bound 1.5e-3, stated in the test)
"""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import oracle_args, rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _das(d, **kw):
    import ffdas_amd
    args = dict(d)
    si = args.pop("sparse_indices", None)
    if si is not None:
        return ffdas_amd.das_sparse(args.pop("x"), args.pop("srcpos"), args.pop("dstpos"),
                                    args.pop("offsets"), args.pop("weights"), si, **args, **kw)
    return ffdas_amd.das(args.pop("x"), args.pop("srcpos"), args.pop("dstpos"),
                         args.pop("offsets"), args.pop("weights"), **args, **kw)


def _ref(w, x=None, **kw):
    a, k = oracle_args(w)
    if x is not None:
        a = (x,) + a[1:]
    k.update(kw)
    return oracle.das(*a, **k)


# ---- reference known-answer tests (bindings/python/tests/test_das.py:83-132) ----------

def _das_args(dev, **overrides):
    import torch
    n_ch, n_seq, n_samp, n_tgt = 4, 2, 16, 8
    args = dict(x=torch.zeros((n_ch, n_seq, n_samp), dtype=torch.float32, device=dev),
                srcpos=torch.zeros((n_ch, 3), dtype=torch.float32, device=dev),
                dstpos=torch.zeros((n_tgt, 3), dtype=torch.float32, device=dev),
                offsets=torch.zeros((n_seq, n_tgt), dtype=torch.float32, device=dev),
                weights=torch.ones((n_seq, n_tgt), dtype=torch.float32, device=dev))
    args.update(overrides)
    return args


def test_zero_input_zero_output(dev):
    import ffdas_amd
    out = ffdas_amd.das(**_das_args(dev))
    assert np.allclose(out.cpu().numpy(), 0.0, atol=1e-6)


def test_zero_weights_zero_output(dev, rng):
    import torch
    import ffdas_amd
    a = _das_args(dev)
    a["x"] = torch.from_numpy(rng.standard_normal(tuple(a["x"].shape)).astype("float32")).to(dev)
    a["weights"] = torch.zeros_like(a["weights"])
    assert np.allclose(ffdas_amd.das(**a).cpu().numpy(), 0.0, atol=1e-6)


def test_output_shape_3d(dev):
    import torch
    import ffdas_amd
    n_ch, n_seq, n_samp, nz, nx = 4, 2, 32, 8, 6
    out = ffdas_amd.das(x=torch.zeros((n_ch, n_seq, n_samp), device=dev),
                        srcpos=torch.zeros((n_ch, 3), device=dev),
                        dstpos=torch.zeros((nz, nx, 3), device=dev),
                        offsets=torch.zeros((n_seq, nz, nx), device=dev),
                        weights=torch.ones((n_seq, nz, nx), device=dev))
    assert tuple(out.shape) == (nz, nx)


def test_output_shape_4d_batched(dev):
    import torch
    import ffdas_amd
    batch, n_ch, n_seq, n_samp, n_tgt = 3, 4, 2, 32, 8
    out = ffdas_amd.das(x=torch.zeros((batch, n_ch, n_seq, n_samp), device=dev),
                        srcpos=torch.zeros((n_ch, 3), device=dev),
                        dstpos=torch.zeros((n_tgt, 3), device=dev),
                        offsets=torch.zeros((n_seq, n_tgt), device=dev),
                        weights=torch.ones((n_seq, n_tgt), device=dev))
    assert tuple(out.shape) == (batch, n_tgt)


@pytest.mark.parametrize("dtype", ["float32", "float64", "float16", "bfloat16"])
def test_single_sample_impulse(dev, dtype):
    """x[0,0,32] = 1, channel and voxel at the origin, offset 32: the radius is clamped to
    sqrt(1e-7) so out = 1 - sqrt(float(1e-7)) exactly (das_alg1_kernels.cuh:64)."""
    import torch
    import ffdas_amd
    x = torch.zeros((1, 1, 64), dtype=getattr(torch, dtype), device=dev)
    x[0, 0, 32] = 1.0
    out = ffdas_amd.das(x=x, srcpos=torch.zeros((1, 3), device=dev),
                        dstpos=torch.zeros((1, 3), device=dev),
                        offsets=torch.full((1, 1), 32.0, device=dev),
                        weights=torch.ones((1, 1), device=dev))
    expect = 1.0 - np.sqrt(np.float64(np.float32(1e-7)))
    assert abs(float(out.squeeze()) - expect) < 1e-6


# ---- parity against the oracle ---------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("srcdir", [False, True])
def test_small_fp32(dev, seed, srcdir):
    w = W.small_random(seed=seed, srcdir=srcdir)
    got = _das(to_dev(w, dev)).cpu().numpy()
    assert rel_l2(got, _ref(w)) < 1e-5


@pytest.mark.parametrize("seed", [0, 3])
def test_small_sparse_fp32(dev, seed):
    w = W.small_random(seed=seed, T=5, K=3, sparse=True)
    got = _das(to_dev(w, dev)).cpu().numpy()
    assert rel_l2(got, _ref(w)) < 1e-5


def test_small_complex_fp32(dev):
    w = W.small_random(seed=2, complex_iq=True)
    got = _das(to_dev(w, dev)).cpu().numpy()
    assert rel_l2(got, _ref(w)) < 1e-5


def test_small_fp64(dev):
    import torch
    w = W.small_random(seed=5, srcdir=True)
    d = to_dev(w, dev, torch.float64)
    got = _das(d).cpu().numpy()
    assert got.dtype == np.float64
    assert rel_l2(got, _ref(w, x=w.x.astype(np.float64))) < 2e-6


def test_config1_fp32(dev):
    w = W.config1()
    got = _das(to_dev(w, dev)).cpu().numpy()
    ref = _ref(w)
    e = rel_l2(got, ref)
    print(f"config1 fp32 rel-L2 vs FP64 = {e:.3e}")
    assert e < 1e-5


@pytest.mark.parametrize("mode", ["use_fp16", "float16_tensor"])
def test_config3_fp16(dev, mode):
    import torch
    w = W.config3()
    ref = _ref(w)
    if mode == "use_fp16":
        got = _das(to_dev(w, dev), use_fp16=True).cpu().numpy()
    else:
        got = _das(to_dev(w, dev, torch.float16)).cpu().numpy()
    e = rel_l2(got, ref)
    eq = rel_l2(got, _ref(w, x=w.x.astype(np.float16).astype(np.float64)))
    print(f"config3 fp16 ({mode}) rel-L2 vs FP64 = {e:.3e}, vs FP64 on fp16 data = {eq:.3e}")
    assert e < 1e-3
    assert eq < 1e-5


def test_config3_bf16(dev):
    """The kernel's own arithmetic is pinned against the oracle run on the bf16-rounded data
    (1e-5, the FP32 bound).  Against the unrounded data the error is the rounding of the data
    itself: bfloat16 keeps 8 significant bits, 2^-9 / sqrt(3) = 1.1e-3 rms relative per
    sample, which the sum over 384 channel-transmit terms carries to about 1e-3 (9.5e-4 on
    this seed), so the bound there is the quantisation floor, 1.5e-3."""
    import torch
    w = W.config3()
    ref = _ref(w)
    got = _das(to_dev(w, dev), use_bf16=True).cpu().numpy()
    xq = torch.from_numpy(w.x).to(torch.bfloat16).to(torch.float64).numpy()
    e = rel_l2(got, ref)
    eq = rel_l2(got, _ref(w, x=xq))
    print(f"config3 bf16 rel-L2 vs FP64 = {e:.3e}, vs FP64 on bf16 data = {eq:.3e}")
    assert eq < 1e-5
    assert e < 1.5e-3


def test_config3_iq_complex(dev):
    import torch
    import ffdas_amd
    w = W.config3(complex_iq=True)
    ref = _ref(w)
    got = _das(to_dev(w, dev)).cpu().numpy()
    e32 = rel_l2(got, ref)
    got16 = _das(to_dev(w, dev, torch.complex32)).cpu().numpy()
    e16 = rel_l2(got16, ref)
    print(f"config3 IQ complex64 rel-L2 = {e32:.3e}, complex32 (half2) rel-L2 = {e16:.3e}")
    assert e32 < 1e-5
    assert e16 < 1e-3


# ---- frame-inner (batched) path: every chunking and storage -----------------------------

@pytest.mark.parametrize("B", [4, 5, 16, 20])
@pytest.mark.parametrize("storage", ["float32", "float16", "bfloat16"])
def test_batched_layout_parity(dev, B, storage):
    """DEFAULT runs the frame-inner kernels for B >= 4 (chunks of 4/8/16 frames, padding
    frames at the tail); ALG1 runs the natural-layout kernel.  Both against the oracle."""
    import torch
    import ffdas_amd
    w = W.small_random(seed=50 + B, B=B, C=7, T=3, S=120, N=(9, 11), srcdir=True)
    xt = getattr(torch, storage)
    d = to_dev(w, dev, xt)
    xq = torch.from_numpy(w.x).to(xt).to(torch.float64).numpy()
    ref = _ref(w, x=xq)
    tol = 1e-5 if storage == "float32" else 1e-4
    for alg in (ffdas_amd.Algorithm.DEFAULT, ffdas_amd.Algorithm.ALG1):
        got = _das(d, algorithm=alg).cpu().numpy()
        assert rel_l2(got, ref) < tol, (alg, rel_l2(got, ref))


@pytest.mark.parametrize("B", [4, 16])
def test_batched_complex_and_sparse(dev, B):
    import torch
    w = W.small_random(seed=70 + B, B=B, C=5, T=4, K=2, S=100, N=(6, 7), complex_iq=True,
                       sparse=True)
    got = _das(to_dev(w, dev)).cpu().numpy()
    assert rel_l2(got, _ref(w)) < 1e-5
    got16 = _das(to_dev(w, dev, torch.complex32)).cpu().numpy()
    xq = torch.from_numpy(w.x).to(torch.complex32).to(torch.complex128).numpy()
    assert rel_l2(got16, _ref(w, x=xq)) < 1e-4


def test_use_fp16_on_batched_float_matches_fp16_tensor(dev):
    import torch
    w = W.small_random(seed=90, B=16, C=6, T=2, S=96, N=(8, 8))
    a = _das(to_dev(w, dev), use_fp16=True).cpu().numpy()
    b = _das(to_dev(w, dev, torch.float16)).cpu().numpy()
    assert np.array_equal(a, b)


def test_channels_trailing_and_strided_input(dev):
    import torch
    w = W.small_random(seed=91, B=8, C=5, T=3, S=80, N=(7, 6))
    d = to_dev(w, dev)
    ref = _das(d).cpu().numpy()
    d2 = dict(d)
    d2["x"] = d["x"].transpose(1, 2).contiguous()
    got = _das(d2, channels_trailing=True).cpu().numpy()
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6)
    big = torch.zeros((8, 5, 3, 100), device=dev)
    big[..., :80] = d["x"]
    d3 = dict(d)
    d3["x"] = big[..., :80]                     # non-contiguous (row stride 100)
    got = _das(d3).cpu().numpy()
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6)


def test_out_and_beta_reuse(dev):
    import torch
    import ffdas_amd
    w = W.small_random(seed=92, B=4)
    d = to_dev(w, dev)
    out = torch.empty((4,) + tuple(w.dstpos.shape[:-1]), device=dev)
    r = _das(d, out=out)
    assert r is out
    assert rel_l2(out.cpu().numpy(), _ref(w)) < 1e-5
