"""The drop-in boundary's helpers on the GPU: Timer (the reference's
bindings/python/tests/test_timer.py:1-39 restated), TensorView on ROCm tensors
(bindings/python/tests/test_types.py:59-76 restated; kDLROCM = 10 where the reference reports
kDLCUDA = 2), a half2 view and a non-torch DLPack producer passed to ``das`` against the FP64
oracle, and ``contiguous_copy`` (ffdas_contiguous_copy) for every type pair of the reference
(src/tensor_ops/contiguous_copy.cu:25-115) on strided, permuted, high-rank and reversed
inputs against torch's own conversions, with the reference's error codes."""
import ctypes
import zlib

import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


# ---- Timer ------------------------------------------------------------------------------

def test_timer_context_manager(dev):
    from ffdas_amd.utils import Timer
    with Timer() as t:
        pass
    ms = t.elapsed_ms()
    assert isinstance(ms, float) and ms >= 0.0


def test_timer_elapsed_before_stop_raises(dev):
    from ffdas_amd.utils import Timer
    with pytest.raises(RuntimeError):
        Timer().elapsed_ms()


def test_timer_double_start_raises(dev):
    from ffdas_amd.utils import Timer
    t = Timer()
    t.start()
    with pytest.raises(RuntimeError):
        t.start()
    t.stop()


def test_timer_manual_start_stop(dev):
    from ffdas_amd.utils import Timer
    t = Timer()
    t.start()
    t.stop()
    assert t.elapsed_ms() >= 0.0


def test_timer_measures_nonzero_work(dev):
    import torch
    from ffdas_amd.utils import Timer
    with Timer() as t:
        a = torch.ones((4096, 4096), dtype=torch.float32, device=dev)
        torch.matmul(a, a)
    assert t.elapsed_ms() > 0.0


# ---- TensorView on device ---------------------------------------------------------------

def test_view_float2_shape(dev):
    import torch
    import ffdas_amd
    v = ffdas_amd.view(torch.zeros((4, 3, 2), dtype=torch.float32, device=dev), ffdas_amd.float2)
    assert v.dtype == ffdas_amd.float2 and v.shape == (4, 3)


def test_view_preserves_base(dev):
    import torch
    import ffdas_amd
    base = torch.ones((8, 2), dtype=torch.float32, device=dev)
    assert ffdas_amd.view(base, ffdas_amd.float2).base is base


def test_view_dlpack_device(dev):
    import torch
    import ffdas_amd
    v = ffdas_amd.view(torch.zeros((4, 2), dtype=torch.float32, device=dev), ffdas_amd.float2)
    dev_type, dev_id = v.__dlpack_device__()
    assert dev_type == 10 and dev_id == (dev.index or 0)       # kDLROCM


def test_view_dlpack_roundtrip_on_device(dev):
    import torch
    import ffdas_amd
    base = torch.randn(64, dtype=torch.complex64, device=dev)
    t = torch.from_dlpack(ffdas_amd.view(base, ffdas_amd.complex64))
    assert t.device == base.device and t.data_ptr() == base.data_ptr()
    assert torch.equal(t, base)


def test_half2_view_into_das(dev):
    """IQ data handed over as a (..., 2) float16 array viewed as half2 (the reference's way
    to pass complex half data, _ffdas.cpp:276-306)."""
    import torch
    import ffdas_amd
    w = W.small_random(seed=31, B=3, C=7, T=3, S=150, N=(9, 8), complex_iq=True)
    pairs = torch.view_as_real(torch.from_numpy(w.x)).half().to(dev)      # (B, C, T, S, 2)
    xv = ffdas_amd.view(pairs, ffdas_amd.half2)
    t = lambda a: torch.from_numpy(a).to(dev)
    out = ffdas_amd.das(xv, t(w.srcpos), t(w.dstpos), t(w.offsets), t(w.weights),
                        wavenum=w.wavenum)
    assert out.dtype == torch.complex64
    xq = torch.view_as_complex(pairs.float().cpu()).numpy().astype(np.complex128)
    ref = oracle.das(xq, w.srcpos, w.dstpos, w.offsets, w.weights, wavenum=w.wavenum)
    assert rel_l2(out.cpu().numpy(), ref) < 2e-5


class _Producer:
    """A DLPack producer that is neither a torch tensor nor a TensorView."""

    def __init__(self, t):
        import ffdas_amd
        self._v = ffdas_amd.view(t, ffdas_amd.float32)

    def __dlpack__(self, stream=None, **kw):
        return self._v.__dlpack__()

    def __dlpack_device__(self):
        return self._v.__dlpack_device__()


def test_foreign_dlpack_producer_into_das(dev):
    import torch
    import ffdas_amd
    w = W.small_random(seed=32, B=2, C=5, T=3, S=128, N=(6, 7))
    t = lambda a: torch.from_numpy(a).to(dev)
    x = t(w.x)
    args = [_Producer(a) for a in (x, t(w.srcpos), t(w.dstpos), t(w.offsets), t(w.weights))]
    out = ffdas_amd.das(*args)
    ref = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


# ---- contiguous_copy ----------------------------------------------------------------------

# (input, output) element types of the reference's dispatch, then the bfloat16 extension
PAIRS = [("int16", "int16"), ("int16", "float32"), ("half", "half"), ("half", "float32"),
         ("float32", "half"), ("float32", "float32"), ("float32", "float64"),
         ("short2", "short2"), ("short2", "float2"), ("half2", "half2"), ("half2", "float2"),
         ("float2", "half2"), ("float2", "float2"), ("float2", "double2"),
         ("float64", "float64"), ("float64", "float32"), ("double2", "double2"),
         ("double2", "float2"),
         ("bfloat16", "bfloat16"), ("bfloat16", "float32"), ("float32", "bfloat16"),
         ("bfloat162", "bfloat162"), ("bfloat162", "float2"), ("float2", "bfloat162")]


def _components(shape, name, dev, seed):
    """A random component tensor for element type `name` (complex pair types get a trailing
    dimension of 2), drawn so that every value is exact in its type."""
    import torch
    import ffdas_amd
    et = getattr(ffdas_amd, name)
    g = torch.Generator(device="cpu").manual_seed(seed)
    full = tuple(shape) + ((2,) if et.pair else ())
    if et.base == torch.int16:
        c = torch.randint(-32768, 32768, full, generator=g, dtype=torch.int32).to(torch.int16)
    else:
        c = (torch.randn(full, generator=g, dtype=torch.float64) * 300.0).to(et.base)
    return c.to(dev), et


def _strided(c, kind, pair):
    """Non-contiguous views of the same logical tensor (pair dim stays last)."""
    import torch
    nd = c.ndim - (1 if pair else 0)
    if kind == "contiguous":
        return c
    if kind == "sliced":
        idx = (slice(None, None, 2),) + (slice(1, None),) * (nd - 1)
        return c[idx]
    if kind == "permuted":
        perm = list(range(nd))[::-1] + ([nd] if pair else [])
        return c.permute(*perm)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["contiguous", "sliced", "permuted"])
@pytest.mark.parametrize("src, dst", PAIRS)
def test_contiguous_copy_pairs(dev, src, dst, kind):
    import torch
    import ffdas_amd
    seed = zlib.crc32(f"{src}{dst}{kind}".encode()) % 1000
    c, et_in = _components((6, 5, 9), src, dev, seed=seed)
    c = _strided(c, kind, et_in.pair)
    et_out = getattr(ffdas_amd, dst)
    xin = ffdas_amd.view(c, et_in)
    shape = xin.shape
    out_base = torch.empty(tuple(shape) + ((2,) if et_out.pair else ()), dtype=et_out.base,
                           device=dev)
    res = ffdas_amd.contiguous_copy(xin, ffdas_amd.view(out_base, et_out))
    assert res.base is out_base
    expect = c.to(et_out.base)                       # torch's conversion: round to nearest even
    assert torch.equal(out_base, expect.contiguous())


@pytest.mark.parametrize("ndim", [1, 5, 7])
def test_contiguous_copy_rank(dev, ndim):
    import torch
    import ffdas_amd
    shape = [3, 2, 4, 1, 3, 2, 5][:ndim]
    g = torch.Generator(device="cpu").manual_seed(ndim)
    big = torch.randn([2 * s + 1 for s in shape], generator=g).to(dev)
    x = big[tuple(slice(1, None, 2) for _ in shape)]
    if ndim > 1:
        x = x.transpose(0, ndim - 1)
    out = ffdas_amd.contiguous_copy(x)
    assert out.is_contiguous() and torch.equal(out, x.contiguous())
    h = ffdas_amd.contiguous_copy(x, torch.empty(tuple(x.shape), dtype=torch.float16,
                                                 device=dev))
    assert torch.equal(h, x.half())


def _desc(dims, strides, dtype):
    from ffdas_amd._lib import TensorDesc
    return TensorDesc(dims, strides, int(dtype))


def _raw_copy(dev, xd, xptr, od, optr):
    import torch
    from ffdas_amd._core import get_library_handle
    from ffdas_amd._lib import lib
    h = get_library_handle(dev.index or 0)
    h.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    rc = lib().ffdas_contiguous_copy(h.h, xd.ptr if xd else None, xptr, od.ptr if od else None,
                                     optr)
    torch.cuda.synchronize(dev)
    return rc


def test_contiguous_copy_negative_strides(dev):
    """A reversed 2-D view: the pointer addresses element [0, 0] (the last in memory)."""
    import torch
    from ffdas_amd._lib import DType
    base = torch.randn((7, 33), device=dev)
    out = torch.empty((7, 33), device=dev)
    last = base.data_ptr() + (base.numel() - 1) * 4
    rc = _raw_copy(dev, _desc([7, 33], [-33, -1], DType.R_32F), ctypes.c_void_p(last),
                   _desc([7, 33], None, DType.R_32F), ctypes.c_void_p(out.data_ptr()))
    assert rc == 0
    assert torch.equal(out, base.flip(0).flip(1))


def test_contiguous_copy_int16_exact(dev):
    import torch
    import ffdas_amd
    x = torch.arange(-32768, 32768, dtype=torch.int32).to(torch.int16).reshape(256, 256).to(dev)
    out = ffdas_amd.contiguous_copy(x.t(), torch.empty((256, 256), device=dev))
    assert torch.equal(out, x.t().float())


def test_contiguous_copy_error_codes(dev):
    import torch
    from ffdas_amd._lib import DType
    x = torch.zeros((4, 5), device=dev)
    y = torch.zeros((4, 5), device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    ok = _desc([4, 5], None, DType.R_32F)
    # different dims, or a non-contiguous output: FFDAS_ERROR_INVALID_DIMS
    assert _raw_copy(dev, ok, p(x), _desc([5, 4], None, DType.R_32F), p(y)) == 2
    assert _raw_copy(dev, ok, p(x), _desc([4, 5], [1, 4], DType.R_32F), p(y)) == 2
    assert _raw_copy(dev, ok, p(x), _desc([4, 5, 1], None, DType.R_32F), p(y)) == 2
    # pairs the reference does not dispatch: FFDAS_ERROR_UNSUPPORTED_TYPE
    for a, b in [(DType.R_16F, DType.R_64F), (DType.R_32I, DType.R_32F),
                 (DType.R_32F, DType.C_32F), (DType.R_16I, DType.R_16F),
                 (DType.C_64F, DType.C_16F), (DType.R_64F, DType.R_16F)]:
        assert _raw_copy(dev, _desc([4, 5], None, a), p(x), _desc([4, 5], None, b), p(y)) == 3
    # missing descriptors and data
    assert _raw_copy(dev, None, p(x), ok, p(y)) == 4
    assert _raw_copy(dev, ok, None, ok, p(y)) == 4
    # nothing to copy
    assert _raw_copy(dev, _desc([0, 5], None, DType.R_32F), None,
                     _desc([0, 5], None, DType.R_32F), None) == 0


# ---- debug mode (synchronous launch check) ------------------------------------------------

def test_debug_mode_waits_for_each_call(dev, monkeypatch):
    """FFDAS_DEBUG=1: every entry point waits for its stream before returning (the reference's
    FFDAS_DEBUG launch check, src/core/error_checking.h:115-126), so the stream is idle right
    after the call; results are those of the asynchronous mode, bit for bit."""
    import torch
    import ffdas_amd
    w = W.config1()
    t = lambda a: torch.from_numpy(a).to(dev)
    args = (t(w.x), t(w.srcpos), t(w.dstpos), t(w.offsets), t(w.weights))
    ref = ffdas_amd.das(*args)
    torch.cuda.synchronize(dev)
    monkeypatch.setenv("FFDAS_DEBUG", "1")
    for _ in range(3):
        out = ffdas_amd.das(*args)
        assert torch.cuda.current_stream(dev).query()
        y = ffdas_amd.contiguous_copy(args[0].transpose(-1, -2))
        assert torch.cuda.current_stream(dev).query()
    assert torch.equal(out, ref)
    assert torch.equal(y, args[0].transpose(-1, -2).contiguous())
