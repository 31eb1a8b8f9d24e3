"""Adjoint parity (GPU): ffdas_das_backward against the FP64 oracle's adjoint, the
dot-product test and autograd plumbing."""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _grads(w, dev, gout, scale=1.0, **kw):
    import torch
    import ffdas_amd
    d = to_dev(w, dev)
    g = torch.from_numpy(gout).to(dev)
    return ffdas_amd.das_backward(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"], g,
                                  srcdir=d.get("srcdir"), wavenum=d.get("wavenum", 0.0),
                                  sparse_indices=d.get("sparse_indices"), scale=scale, **kw)


@pytest.mark.parametrize("case", ["dense", "dir", "sparse", "complex", "scaled"])
def test_small_adjoint_fp32(dev, case):
    kw = {}
    scale = 1.0
    if case == "dense":
        w = W.small_random(seed=11)
    elif case == "dir":
        w = W.small_random(seed=12, srcdir=True)
    elif case == "sparse":
        w = W.small_random(seed=13, T=5, K=3, sparse=True)
    elif case == "complex":
        w = W.small_random(seed=14, complex_iq=True)
    else:
        w = W.small_random(seed=15)
        scale = 1.03
    rng = np.random.default_rng(7)
    B = w.x.shape[0]
    shape = (B,) + w.dstpos.shape[:-1]
    gout = rng.standard_normal(shape).astype(np.float32)
    if np.iscomplexobj(w.x):
        gout = (gout + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices,
                             scale=scale)
    gx, go, gw, gs = _grads(w, dev, gout, scale=scale)
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), ref["grad_offsets"]) < 1e-4
    assert rel_l2(gw.cpu().numpy(), ref["grad_weights"]) < 1e-5
    assert abs(float(gs) - ref["grad_scale"]) <= 1e-4 * max(1.0, abs(ref["grad_scale"]))


def test_dot_product_config1_fp32(dev):
    """<A x, y> == <x, A^T y> with the library's own forward and adjoint."""
    import torch
    import ffdas_amd
    w = W.config1()
    d = to_dev(w, dev)
    y = torch.randn((1,) + tuple(w.dstpos.shape[:-1]), device=dev, generator=None)
    ax = ffdas_amd.das(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    gx, _, _, _ = ffdas_amd.das_backward(d["x"], d["srcpos"], d["dstpos"], d["offsets"],
                                         d["weights"], y, need_offsets=False,
                                         need_weights=False, need_scale=False)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((d["x"].double() * gx.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))


def test_autograd_matches_backward(dev):
    import torch
    import ffdas_amd
    w = W.small_random(seed=21, B=3)
    d = to_dev(w, dev)
    x = d["x"].clone().requires_grad_(True)
    ofs = d["offsets"].clone().requires_grad_(True)
    s = torch.tensor(1.01, device=dev, requires_grad=True)
    out = ffdas_amd.das(x, d["srcpos"], d["dstpos"], ofs, d["weights"], scale=s)
    g = torch.randn_like(out)
    (out * g).sum().backward()
    gx, go, gw, gs = ffdas_amd.das_backward(d["x"], d["srcpos"], d["dstpos"], d["offsets"],
                                            d["weights"], g, scale=1.01)
    assert torch.allclose(x.grad, gx, rtol=1e-6, atol=1e-6)
    assert torch.allclose(ofs.grad, go, rtol=1e-6, atol=1e-6)
    assert abs(float(s.grad) - float(gs)) <= 1e-4 * max(1.0, abs(float(gs)))


@pytest.mark.parametrize("B", [5, 16])
@pytest.mark.parametrize("storage", ["float32", "float16"])
def test_batched_adjoint(dev, B, storage):
    """Frame-inner parameter kernel and chunked grad_x scatter (B > frames per chunk)."""
    import torch
    import ffdas_amd
    w = W.small_random(seed=30 + B, B=B, C=6, T=3, S=110, N=(8, 9), srcdir=True)
    rng = np.random.default_rng(8)
    gout = rng.standard_normal((B,) + w.dstpos.shape[:-1]).astype(np.float32)
    xt = getattr(torch, storage)
    xq = torch.from_numpy(w.x).to(xt).to(torch.float64).numpy()
    ref = oracle.das_adjoint(xq, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             srcdir=w.srcdir, scale=0.98)
    d = to_dev(w, dev, xt)
    gx, go, gw, gs = ffdas_amd.das_backward(d["x"], d["srcpos"], d["dstpos"], d["offsets"],
                                            d["weights"], torch.from_numpy(gout).to(dev),
                                            srcdir=d["srcdir"], scale=0.98)
    tol = 1e-5 if storage == "float32" else 5e-4
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), ref["grad_offsets"]) < max(tol, 1e-4)
    assert rel_l2(gw.cpu().numpy(), ref["grad_weights"]) < max(tol, 1e-5)
    assert abs(float(gs) - ref["grad_scale"]) <= max(tol, 1e-4) * max(1.0, abs(ref["grad_scale"]))


def test_grad_x_fixed_point_path_is_deterministic(dev, monkeypatch):
    """The default grad_x kernel accumulates in 64-bit fixed point: repeated adjoints are
    bitwise identical.  FFDAS_BWDX=sorted selects the bucketed FP32 kernel, which agrees to
    float rounding."""
    import torch
    import ffdas_amd
    w = W.config1()
    d = to_dev(w, dev)
    g = torch.randn((1,) + tuple(w.dstpos.shape[:-1]), device=dev)
    args = (d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"], g)
    kw = dict(need_offsets=False, need_weights=False, need_scale=False)
    a = ffdas_amd.das_backward(*args, **kw)[0]
    b = ffdas_amd.das_backward(*args, **kw)[0]
    assert torch.equal(a, b)
    monkeypatch.setenv("FFDAS_BWDX", "sorted")
    sorted_gx = ffdas_amd.das_backward(*args, **kw)[0]
    assert rel_l2(sorted_gx.cpu().numpy(), a.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("S", [700, 1500, 4000, 5000])
@pytest.mark.parametrize("cplx", [False, True])
def test_grad_x_row_lengths(dev, S, cplx):
    """Rows of several lengths (LDS row sizes, frames per workgroup) for the grad_x kernel."""
    import torch
    import ffdas_amd
    w = W.small_random(seed=S, B=3, C=3, T=2, S=S, N=(12, 10), complex_iq=cplx)
    w.offsets[:] = np.random.default_rng(S).uniform(-10, S * 0.9, w.offsets.shape).astype(np.float32)
    rng = np.random.default_rng(9)
    gout = rng.standard_normal((3,) + w.dstpos.shape[:-1])
    if cplx:
        gout = gout + 1j * rng.standard_normal(gout.shape)
    gout = gout.astype(np.complex64 if cplx else np.float32)
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             wavenum=w.wavenum)
    gx, _, _, _ = _grads(w, dev, gout, need_offsets=False, need_weights=False, need_scale=False)
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5


@pytest.mark.parametrize("N", [(5,), (70, 3), (3, 4, 5), (1, 40, 2), (130, 1, 3), (1, 1)])
def test_grad_x_grid_shapes(dev, N):
    """Axial voxel order of the grad_x scatter: first dimension shorter and longer than a
    wave, unit dimensions, 1-D and 3-D grids, a single voxel."""
    w = W.small_random(seed=sum(N), B=3, C=4, T=2, S=120, N=N)
    rng = np.random.default_rng(5)
    gout = rng.standard_normal((3,) + w.dstpos.shape[:-1]).astype(np.float32)
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, gout)
    gx, _, _, _ = _grads(w, dev, gout, need_offsets=False, need_weights=False, need_scale=False)
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5


@pytest.mark.parametrize("B, cplx", [(16, False), (24, False), (20, False), (8, True)])
def test_grad_x_last_round_split(dev, monkeypatch, B, cplx):
    """More rows than one round of workgroups (320 rows on 256 CUs): the rows of the last,
    partly filled round run in workgroups of half the frames, reading half records of the
    transposed grad_out (24 and 20 frames: two record groups, the last one partial).  The
    fixed-point sums are the same integers, so grad_x is bitwise equal to the unsplit launch
    (FFDAS_BWDX_TAIL=0), and it matches the oracle."""
    import torch
    w = W.small_random(seed=B, B=B, C=64, T=5, S=400, N=(16, 8), complex_iq=cplx)
    w.offsets[:] = np.random.default_rng(B).uniform(-10, 360, w.offsets.shape).astype(np.float32)
    rng = np.random.default_rng(B + 1)
    gout = rng.standard_normal((B,) + w.dstpos.shape[:-1])
    if cplx:
        gout = gout + 1j * rng.standard_normal(gout.shape)
    gout = gout.astype(np.complex64 if cplx else np.float32)
    kw = dict(need_offsets=False, need_weights=False, need_scale=False)
    split = _grads(w, dev, gout, **kw)[0]
    monkeypatch.setenv("FFDAS_BWDX_TAIL", "0")
    whole = _grads(w, dev, gout, **kw)[0]
    assert torch.equal(split, whole)
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             wavenum=w.wavenum)
    assert rel_l2(split.cpu().numpy(), ref["grad_x"]) < 1e-5


@pytest.mark.parametrize("cplx, fp16", [(False, True), (True, True), (False, False)])
def test_concurrent_adjoint_matches_serial(dev, monkeypatch, cplx, fp16):
    """grad_x on the handle's second stream beside the parameter adjoint (the default for
    16-bit storage when both are asked for; forced with FFDAS_BWD_CONCURRENT=1 for FP32)
    gives bitwise the results of the two run one after the other (FFDAS_BWD_CONCURRENT=0),
    and matches the oracle's adjoint (FP32 storage) or its storage-rounded data (FP16)."""
    import torch
    w = W.small_random(seed=77, B=16, C=24, T=3, S=600, N=(40, 24), complex_iq=cplx)
    w.offsets[:] = np.random.default_rng(7).uniform(-10, 520, w.offsets.shape).astype(np.float32)
    rng = np.random.default_rng(8)
    gout = rng.standard_normal((16,) + w.dstpos.shape[:-1])
    if cplx:
        gout = gout + 1j * rng.standard_normal(gout.shape)
    gout = gout.astype(np.complex64 if cplx else np.float32)
    if not fp16:
        monkeypatch.setenv("FFDAS_BWD_CONCURRENT", "1")
    both = _grads(w, dev, gout, use_fp16=fp16)
    monkeypatch.setenv("FFDAS_BWD_CONCURRENT", "0")
    serial = _grads(w, dev, gout, use_fp16=fp16)
    for a, b in zip(both, serial):
        assert torch.equal(a, b)
    xr = w.x
    if fp16:   # the kernel gathers the FP16-rounded data
        xr = (w.x.real.astype(np.float16).astype(np.float32) +
              1j * w.x.imag.astype(np.float16).astype(np.float32)).astype(np.complex64) \
            if cplx else w.x.astype(np.float16).astype(np.float32)
    ref = oracle.das_adjoint(xr, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             wavenum=w.wavenum)
    tol = 5e-4 if fp16 else 1e-5
    assert rel_l2(both[0].cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(both[1].cpu().numpy(), ref["grad_offsets"]) < max(tol, 1e-4)
    assert rel_l2(both[2].cpu().numpy(), ref["grad_weights"]) < tol


@pytest.mark.parametrize("case", ["dense", "dir", "sparse", "complex", "scaled", "float64", "batched"])
def test_plain_adjoint_alg1(dev, case):
    """ALG1 runs the plain kernels: natural-layout parameter gather and the global-atomic
    grad_x scatter; parity with the FP64 oracle's adjoint as for the tuned path."""
    import torch
    import ffdas_amd
    scale = 1.0
    xt = None
    if case == "dense":
        w = W.small_random(seed=41)
    elif case == "dir":
        w = W.small_random(seed=42, srcdir=True)
    elif case == "sparse":
        w = W.small_random(seed=43, T=5, K=3, sparse=True)
    elif case == "complex":
        w = W.small_random(seed=44, complex_iq=True)
    elif case == "scaled":
        w = W.small_random(seed=45)
        scale = 0.97
    elif case == "float64":
        w = W.small_random(seed=46)
        xt = torch.float64
    else:
        w = W.small_random(seed=47, B=20, C=5, T=3, S=100, N=(9, 8))
    rng = np.random.default_rng(9)
    shape = (w.x.shape[0],) + w.dstpos.shape[:-1]
    gout = rng.standard_normal(shape)
    if np.iscomplexobj(w.x):
        gout = gout + 1j * rng.standard_normal(shape)
    gout = gout.astype(np.complex64 if np.iscomplexobj(w.x) else (np.float64 if xt else np.float32))
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices,
                             scale=scale)
    d = to_dev(w, dev, xt)
    gx, go, gw, gs = ffdas_amd.das_backward(
        d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"], torch.from_numpy(gout).to(dev),
        srcdir=d.get("srcdir"), wavenum=d.get("wavenum", 0.0),
        sparse_indices=d.get("sparse_indices"), scale=scale, algorithm=ffdas_amd.Algorithm.ALG1)
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), ref["grad_offsets"]) < 1e-4
    assert rel_l2(gw.cpu().numpy(), ref["grad_weights"]) < 1e-5
    assert abs(float(gs) - ref["grad_scale"]) <= 1e-4 * max(1.0, abs(ref["grad_scale"]))


def test_reuse_x_matches_a_fresh_copy(dev):
    """reuse_x gathers from the forward's converted copy: identical gradients when x is
    unchanged; a forward of another x in between is detected and a fresh copy is made."""
    import torch
    import ffdas_amd
    w = W.small_random(seed=61, B=16, C=6, T=3, S=120, N=(12, 10))
    w2 = W.small_random(seed=62, B=16, C=6, T=3, S=120, N=(12, 10))
    d, d2 = to_dev(w, dev, torch.float16), to_dev(w2, dev, torch.float16)
    args = lambda q: (q["x"], q["srcpos"], q["dstpos"], q["offsets"], q["weights"])
    g = torch.randn((16, 12, 10), device=dev)
    fresh = ffdas_amd.das_backward(*args(d), g)
    ffdas_amd.das(*args(d))
    reused = ffdas_amd.das_backward(*args(d), g, reuse_x=True)
    ffdas_amd.das(*args(d))
    ffdas_amd.das(*args(d2))                     # another x overwrites the copy
    other = ffdas_amd.das_backward(*args(d), g, reuse_x=True)
    for a, b, c in zip(fresh, reused, other):
        assert torch.equal(a, b)
        assert torch.equal(a, c)


@pytest.mark.parametrize("K,T,N", [(3, 5, (9, 8)), (1, 4, (9, 8)), (6, 2, (9, 8)),
                                   (3, 300, (40, 60)), (2, 4100, (9, 8))])
def test_sparse_adjoint_inverted_index(dev, K, T, N):
    """The sparse grad_x visits each transmit's entries through an inverted index (block-
    ordered for T <= 4096, several 1024-voxel blocks at 40 x 60; atomically filled above);
    indices outside [0, T) select nothing, in the forward and in every gradient."""
    w = W.small_random(seed=70 + K, B=5, C=4, T=T, K=K, S=110, N=N, sparse=True)
    bad = np.random.default_rng(K).random(w.sparse_indices.shape) < 0.15
    w.sparse_indices[bad] = np.where(np.random.default_rng(1).random(bad.sum()) < 0.5, -1, T)
    rng = np.random.default_rng(11)
    gout = rng.standard_normal((5,) + w.dstpos.shape[:-1]).astype(np.float32)
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             sparse_indices=w.sparse_indices)
    gx, go, gw, gs = _grads(w, dev, gout)
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), ref["grad_offsets"]) < 1e-4
    assert rel_l2(gw.cpu().numpy(), ref["grad_weights"]) < 1e-5
    assert abs(float(gs) - ref["grad_scale"]) <= 1e-4 * max(1.0, abs(ref["grad_scale"]))


@pytest.mark.parametrize("B", [8, 16])
@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_sparse_frame_rows_16bit(dev, B, storage):
    """Sparse compounding gathers its frame-inner copy in whole frame rows (dense problems
    in frame-group planes): 16-bit storage, chunks of 8 and 16 frames, with indices outside
    [0, T) that select nothing, forward and every gradient against the oracle on the same
    rounded channel data."""
    import torch
    import ffdas_amd
    w = W.small_random(seed=300 + B, B=B, C=5, T=6, K=3, S=130, N=(10, 9), sparse=True,
                       srcdir=True)
    bad = np.random.default_rng(B).random(w.sparse_indices.shape) < 0.1
    w.sparse_indices[bad] = -1
    xt = getattr(torch, storage)
    xq = torch.from_numpy(w.x).to(xt).to(torch.float64).numpy()
    d = to_dev(w, dev, xt)
    out = ffdas_amd.das_sparse(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"],
                               d["sparse_indices"], srcdir=d["srcdir"]).cpu().numpy()
    ref_out = oracle.das(xq, w.srcpos, w.dstpos, w.offsets, w.weights,
                         sparse_indices=w.sparse_indices, srcdir=w.srcdir)
    assert rel_l2(out, ref_out) < 2e-5
    rng = np.random.default_rng(12)
    gout = rng.standard_normal((B,) + w.dstpos.shape[:-1]).astype(np.float32)
    ref = oracle.das_adjoint(xq, w.srcpos, w.dstpos, w.offsets, w.weights, gout,
                             sparse_indices=w.sparse_indices, srcdir=w.srcdir)
    gx, go, gw, gs = ffdas_amd.das_backward(d["x"], d["srcpos"], d["dstpos"], d["offsets"],
                                            d["weights"], torch.from_numpy(gout).to(dev),
                                            sparse_indices=d["sparse_indices"],
                                            srcdir=d["srcdir"])
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), ref["grad_offsets"]) < 1e-4
    assert rel_l2(gw.cpu().numpy(), ref["grad_weights"]) < 2e-5
    assert abs(float(gs) - ref["grad_scale"]) <= 1e-4 * max(1.0, abs(ref["grad_scale"]))
