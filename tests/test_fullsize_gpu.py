"""Parity at the benchmark configurations' full sizes (GPU).  The whole output is computed on
the GPU; the FP64 oracle evaluates a seeded sample of voxels (all frames, channels and
transmits of each), and size-independent properties are checked over everything: linearity
of the forward and the adjoint identity <A x, y> = <x, A^T y>."""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _sample(w, n, seed=0):
    """Random voxel subset: (flat indices, dstpos (n,3), offsets/weights/sparse (K,n))."""
    N = int(np.prod(w.dstpos.shape[:-1]))
    idx = np.sort(np.random.default_rng(seed).choice(N, size=min(n, N), replace=False))
    K = w.offsets.shape[0]
    dst = w.dstpos.reshape(-1, 3)[idx]
    ofs = w.offsets.reshape(K, -1)[:, idx]
    wts = w.weights.reshape(K, -1)[:, idx]
    sp = None if w.sparse_indices is None else w.sparse_indices.reshape(K, -1)[:, idx]
    return idx, dst, ofs, wts, sp


def _sampled_ref(w, x, n, seed=0):
    idx, dst, ofs, wts, sp = _sample(w, n, seed)
    ref = oracle.das(x, w.srcpos, dst, ofs, wts, srcdir=w.srcdir, wavenum=w.wavenum,
                     sparse_indices=sp)
    return idx, ref


@pytest.fixture(scope="module")
def c4():
    return W.config4()


def test_config4_full_size_fp32(dev, c4):
    import torch
    import ffdas_amd
    d = to_dev(c4, dev)
    args = (d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    out = ffdas_amd.das(d["x"], *args)
    idx, ref = _sampled_ref(c4, c4.x, 400)
    got = out.reshape(out.shape[0], -1)[:, torch.from_numpy(idx).to(dev)].cpu().numpy()
    assert rel_l2(got, ref) < 1e-5
    # linearity over the whole output
    x2 = torch.roll(d["x"], shifts=37, dims=-1)
    lhs = ffdas_amd.das(d["x"] + 2.0 * x2, *args)
    rhs = out + 2.0 * ffdas_amd.das(x2, *args)
    assert float((lhs - rhs).norm() / rhs.norm()) < 1e-6


def test_config4_full_size_fp16(dev, c4):
    import torch
    import ffdas_amd
    d = to_dev(c4, dev, torch.float16)
    out = ffdas_amd.das(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    xq = torch.from_numpy(c4.x).half().double().numpy()
    idx, ref_q = _sampled_ref(c4, xq, 400, seed=1)
    _, ref = _sampled_ref(c4, c4.x, 400, seed=1)
    got = out.reshape(out.shape[0], -1)[:, torch.from_numpy(idx).to(dev)].cpu().numpy()
    assert rel_l2(got, ref_q) < 1e-4
    assert rel_l2(got, ref) < 1e-3


def test_config4_full_size_adjoint(dev, c4):
    """Dot-product identity over all 16 frames, and grad_x of frame 0 against the oracle."""
    import torch
    import ffdas_amd
    d = to_dev(c4, dev)
    args = (d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    g = torch.Generator(device="cpu").manual_seed(7)
    y = torch.randn((16, 256, 256), generator=g).to(dev)
    ax = ffdas_amd.das(d["x"], *args)
    gx = ffdas_amd.das_backward(d["x"], *args, y, need_offsets=False, need_weights=False,
                                need_scale=False)[0]
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((d["x"].double() * gx.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))
    ref = oracle.das_adjoint(c4.x[:1], c4.srcpos, c4.dstpos, c4.offsets, c4.weights,
                             y[:1].cpu().numpy())["grad_x"]
    assert rel_l2(gx[:1].cpu().numpy(), ref) < 1e-5


def test_config2_full_size_3d(dev):
    """The volumetric configuration: 1024 channels, 16 diverging waves, 4096 samples, 128^3
    voxels, directivity mask and Hann apodization (3.4e10 units in one call)."""
    import torch
    import ffdas_amd
    w = W.config2()
    d = to_dev(w, dev)
    out = ffdas_amd.das(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"],
                        srcdir=d["srcdir"])
    idx, ref = _sampled_ref(w, w.x, 300, seed=2)
    got = out.reshape(out.shape[0], -1)[:, torch.from_numpy(idx).to(dev)].cpu().numpy()
    assert rel_l2(got, ref) < 1e-5


def test_config5_full_geometry(dev):
    """The three sparse-compounding probes at their full grids and channel counts (two frames
    of the ensemble), in one batched FP16 call."""
    import torch
    import ffdas_amd
    rng = np.random.default_rng(5)
    probs, ws = [], []
    for name in W.CONFIG5_PROBES:
        elem, grid, offs, wts, sidx, _ = W.config5_geometry(name)
        C = elem.shape[0]
        x = rng.standard_normal((2, C, 64, 2048), dtype=np.float32)
        x = x.astype(np.float16).astype(np.float32)            # exactly representable
        w = W.Workload(name, x, (elem * W.KS).astype(np.float32),
                       (grid * W.KS).astype(np.float32), offs, wts, sparse_indices=sidx)
        ws.append(w)
        d = to_dev(w, dev, torch.float16)
        probs.append(dict(x=d["x"], srcpos=d["srcpos"], dstpos=d["dstpos"],
                          offsets=d["offsets"], weights=d["weights"],
                          sparse_indices=d["sparse_indices"]))
    outs = ffdas_amd.das_batched(probs)
    for i, (w, o) in enumerate(zip(ws, outs)):
        idx, ref = _sampled_ref(w, w.x, 128, seed=10 + i)
        got = o.reshape(o.shape[0], -1)[:, torch.from_numpy(idx).to(dev)].cpu().numpy()
        # The data are exactly representable in FP16 and every kernel uses the compensated
        # FP32 delays, so the FP32-accumulation bound of the FP32 path applies.
        assert rel_l2(got, ref) < 1e-5, w.name


def _storage_tensor(x, storage):
    import torch
    return {"fp16": x.half(), "bf16": x.bfloat16(), "fp32": x}[storage]


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_config4_full_size_adjoint_16bit(dev, c4, storage):
    """The adjoint the benchmark times, at its full size: 16-bit channel data, all 16 frames,
    delay scale 1.02; grad_x, grad_offsets, grad_weights and grad_scale against the FP64
    oracle adjoint of the same (storage-rounded) data.

    The scale handed to the oracle is the FP32 value the kernels receive.  Bounds (rel-L2;
    grad_scale relative), those of the FP32 adjoint tests: grad_x and grad_weights 1e-5,
    grad_offsets and grad_scale 1e-4 (every kernel uses the compensated FP32 delays; the data
    are compared as stored)."""
    import torch
    import ffdas_amd
    d = to_dev(c4, dev)
    x = _storage_tensor(d["x"], storage)
    args = (d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    g = torch.Generator(device="cpu").manual_seed(7)
    y = torch.randn((16, 256, 256), generator=g)
    s = float(np.float32(1.02))
    gx, go, gw, gs = ffdas_amd.das_backward(x, *args, y.to(dev), scale=s)
    xq = x.double().cpu().numpy()
    ref = oracle.das_adjoint(xq, c4.srcpos, c4.dstpos, c4.offsets, c4.weights, y.numpy(), scale=s)
    e_x = rel_l2(gx.cpu().numpy(), ref["grad_x"])
    e_o = rel_l2(go.cpu().numpy(), ref["grad_offsets"])
    e_w = rel_l2(gw.cpu().numpy(), ref["grad_weights"])
    e_s = abs(float(gs) - ref["grad_scale"]) / abs(ref["grad_scale"])
    print(f"{storage}: grad_x {e_x:.2e} grad_offsets {e_o:.2e} grad_weights {e_w:.2e} "
          f"grad_scale {e_s:.2e}")
    assert e_x < 1e-5 and e_w < 1e-5
    assert e_o < 1e-4 and e_s < 1e-4


def _loss(out, target):
    return float(((out.double() - target) ** 2).sum())


def test_config4_scale_finite_difference(dev, c4):
    """Central finite differences of the L2 loss with respect to the delay scale, h = 1e-3, on
    the benchmark path (FP16 channel data, all 16 frames, s = 1.02).

    At h = 1e-3 the delays move by about 1.5 samples (a fifth of a 5 MHz period at 40 MHz), so
    the h = 1e-3 quotient differs from the derivative by tens of per cent through curvature
    alone.  Three checks therefore: (1) the GPU quotient at h = 1e-3 equals the FP64 oracle's
    quotient at h = 1e-3 (the same finite-difference quantity; checks the forward at s +- h);
    (2) the GPU grad_scale equals the FP64 oracle adjoint's; (3) the GPU quotient converges to
    the GPU grad_scale: at h = 1e-5 within 3e-3 (truncation ~7e-4 plus FP32 output rounding)."""
    import torch
    import ffdas_amd
    d = to_dev(c4, dev)
    x = d["x"].half()
    args = (d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    g = torch.Generator(device="cpu").manual_seed(11)
    target = torch.randn((16, 256, 256), generator=g, dtype=torch.float64).to(dev)
    s0 = float(np.float32(1.02))
    f32 = lambda v: float(np.float32(v))        # the scale the kernels receive
    L = lambda s: _loss(ffdas_amd.das(x, *args, scale=s), target)
    out = ffdas_amd.das(x, *args, scale=s0)
    gs = ffdas_amd.das_backward(x, *args, (2.0 * (out.double() - target)).float(), scale=s0,
                                need_x=False, need_offsets=False, need_weights=False)[3]
    gs = float(gs)
    xq = x.double().cpu().numpy()
    tq = target.cpu().numpy()
    Lo = lambda s: float(((oracle.das(xq, c4.srcpos, c4.dstpos, c4.offsets, c4.weights,
                                       scale=f32(s)) - tq) ** 2).sum())
    h = 1e-3
    fd_gpu = (L(s0 + h) - L(s0 - h)) / (f32(s0 + h) - f32(s0 - h))
    fd_ref = (Lo(s0 + h) - Lo(s0 - h)) / (f32(s0 + h) - f32(s0 - h))
    o0 = oracle.das(xq, c4.srcpos, c4.dstpos, c4.offsets, c4.weights, scale=s0)
    gs_ref = oracle.das_adjoint(xq, c4.srcpos, c4.dstpos, c4.offsets, c4.weights,
                                2.0 * (o0 - tq), scale=s0)["grad_scale"]
    fd_small = (L(s0 + 1e-5) - L(s0 - 1e-5)) / (f32(s0 + 1e-5) - f32(s0 - 1e-5))
    print(f"grad_scale gpu {gs:.6e} oracle {gs_ref:.6e}; FD(1e-3) gpu {fd_gpu:.6e} "
          f"oracle {fd_ref:.6e}; FD(1e-5) gpu {fd_small:.6e}")
    assert abs(fd_gpu - fd_ref) <= 1e-4 * abs(fd_ref)
    assert abs(gs - gs_ref) <= 1e-4 * abs(gs_ref)
    assert abs(fd_small - gs) <= 3e-3 * abs(gs)


def test_config4_offsets_finite_difference(dev, c4):
    """Central finite differences (h = 1e-3 samples) of the L2 loss with respect to single
    transmit delays on the benchmark path (FP16 data, 16 frames, s = 1.02), against grad_offsets.
    offsets[j, n] only moves voxel n, so the loss difference is taken over that voxel's 16
    outputs, and the quotient divides by the actual FP32 step.  Interpolation is piecewise
    linear in the delay, so the quotient is exact unless a sample boundary lies within h: such
    an entry is checked against the FP64 oracle's quotient for the same perturbed delays
    instead.  The bound 2e-3 covers FP32 output rounding."""
    import torch
    import ffdas_amd
    d = to_dev(c4, dev)
    x = d["x"].half()
    g = torch.Generator(device="cpu").manual_seed(12)
    target = torch.randn((16, 256, 256), generator=g).to(dev)
    s0 = float(np.float32(1.02))
    ofs = d["offsets"]
    args = lambda o: (d["srcpos"], d["dstpos"], o, d["weights"])
    out = ffdas_amd.das(x, *args(ofs), scale=s0)
    go = ffdas_amd.das_backward(x, *args(ofs), 2.0 * (out - target), scale=s0, need_x=False,
                                need_weights=False, need_scale=False)[1]
    rng = np.random.default_rng(5)
    h = 1e-3
    checked = 0
    for _ in range(40):
        j, zi, xi = int(rng.integers(11)), int(rng.integers(256)), int(rng.integers(256))
        vals, pert = [], []
        for sign in (1.0, -1.0):
            o = ofs.clone()
            o[j, zi, xi] += sign * h            # rounded to FP32: divide by the actual step
            pert.append(float(o[j, zi, xi]))
            y = ffdas_amd.das(x, *args(o), scale=s0)[:, zi, xi].double()
            vals.append(float(((y - target[:, zi, xi].double()) ** 2).sum()))
        fd = (vals[0] - vals[1]) / (pert[0] - pert[1])
        gv = float(go[j, zi, xi])
        scale_ref = max(abs(gv), 1e-3 * float(go.abs().max()))
        if abs(fd - gv) > 2e-3 * scale_ref:
            # a sample boundary within h of this delay changes the slope: confirm with the
            # oracle's quotient before calling it a mismatch
            xq = x[:, :, :, :].double().cpu().numpy()
            dst = c4.dstpos[zi:zi + 1, xi:xi + 1]
            o_np = c4.offsets[:, zi:zi + 1, xi:xi + 1].astype(np.float64)
            w_np = c4.weights[:, zi:zi + 1, xi:xi + 1]
            tq = target[:, zi, xi].double().cpu().numpy()
            ls = []
            for v in pert:
                oo = o_np.copy()
                oo[j] = v
                r = oracle.das(xq, c4.srcpos, dst, oo.astype(np.float32), w_np, scale=s0)
                ls.append(float(((r[:, 0, 0] - tq) ** 2).sum()))
            fd_ref = (ls[0] - ls[1]) / (pert[0] - pert[1])
            assert abs(fd - fd_ref) <= 2e-3 * max(abs(fd_ref), scale_ref), (j, zi, xi, fd, fd_ref)
            continue
        checked += 1
    assert checked >= 30
