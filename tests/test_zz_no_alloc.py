"""ffdas_das_noalloc through the C-ABI (GPU): the forward whose converted copy of the channel
data lives in the library's static device array instead of the handle's workspace.  Its
result must equal ffdas_das's on the same inputs (the same kernels: bit for bit) on two calls
in a row, the second with a smaller input; a problem whose copy exceeds the array is refused
with FFDAS_ERROR_DIMS_TOO_LARGE."""
import ctypes

import numpy as np
import pytest

from ffdas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _call(name, dev, w, x, compute, alg=0):
    import torch
    from ffdas_amd._core import element_code, get_library_handle
    from ffdas_amd._lib import TensorDesc, lib
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    srcpos, dstpos, ofs, wts = t(w.srcpos), t(w.dstpos), t(w.offsets), t(w.weights)
    out = torch.full((x.shape[0],) + tuple(w.dstpos.shape[:-1]), float("nan"),
                     dtype=torch.float32, device=dev)
    xd = TensorDesc(tuple(x.shape), tuple(x.stride()), element_code(x))
    od = TensorDesc(tuple(out.shape), tuple(out.stride()), element_code(out))
    beta = (ctypes.c_float * 1)(0.0)
    p = lambda a: ctypes.c_void_p(a.data_ptr())
    h = get_library_handle(dev.index or 0)
    h.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    rc = getattr(lib(), name)(h.h, p(srcpos), None, 0.0, xd.ptr, p(x), p(dstpos), p(ofs), p(wts),
                              ctypes.cast(beta, ctypes.c_void_p), od.ptr, p(out), int(compute),
                              int(alg))
    torch.cuda.synchronize(dev)
    return rc, out


def test_noalloc_matches_das_two_calls(dev):
    import torch
    from ffdas_amd._lib import ComputeType
    print("case: config-4 frames, FP16 storage, then a smaller FP32 problem", flush=True)
    big = W.config4(n_frames=16)
    small = W.small_random(seed=41, B=5, C=9, T=4, S=300, N=(13, 17))
    for w, ct in ((big, ComputeType.FP16), (small, ComputeType.FP32)):
        x = torch.from_numpy(w.x).to(dev)
        rc0, ref = _call("ffdas_das", dev, w, x, ct)
        rc1, got = _call("ffdas_das_noalloc", dev, w, x, ct)
        assert rc0 == 0 and rc1 == 0
        assert torch.equal(got, ref)


def test_noalloc_refuses_a_copy_beyond_the_array(dev):
    import torch
    from ffdas_amd._lib import ComputeType
    # 16 frames x 256 channels x 16 transmits x 4096 samples in FP32: a 1 GiB copy
    w = W.small_random(seed=42, B=1, C=256, T=16, S=4096, N=(4, 4))
    x = torch.zeros((16, 256, 16, 4096), dtype=torch.float32, device=dev)
    rc, _ = _call("ffdas_das_noalloc", dev, w, x, ComputeType.FP32)
    assert rc == 7                                      # FFDAS_ERROR_DIMS_TOO_LARGE
    rc, ref = _call("ffdas_das", dev, w, x, ComputeType.FP32)
    assert rc == 0 and bool(torch.isfinite(ref).all())
