"""CPU restatement of ffdas's `greens` (frequency-domain Green's function summation).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU leg as the checker; the
library never calls it.

out[b, t, f] = sum_s x[b, s, f] exp(i k_f r_ts) / r_ts with r_ts = |dst_t - src_s|
(src/greens/greens_kernels.cuh:43-73 per term; docs/api/greens.md), in FP64.

Pinning (tests/test_greens.py): the reference's own test formula and cases
(bindings/python/tests/test_greens.py:8-63), a single source (a closed form: one spherical
wave), linearity in x, and the far-field limit (the sum over a small source patch seen from
far away approaches the patch's total weight times e^{ikR}/R)."""
import numpy as np


def greens(srcpos, wavenums, x, dstpos):
    """x: ([batch,] sources, frequencies); returns ([batch,] dst..., frequencies) complex128."""
    src = np.asarray(srcpos, np.float64).reshape(-1, 3)
    dst = np.asarray(dstpos, np.float64)
    k = np.asarray(wavenums, np.float64)
    xx = np.asarray(x)
    squeeze = xx.ndim == 2
    if squeeze:
        xx = xx[None]
    flat = dst.reshape(-1, 3)
    r = np.sqrt(((flat[:, None, :] - src[None, :, :]) ** 2).sum(-1))      # (t, s)
    g = np.exp(1j * k[None, None, :] * r[:, :, None]) / r[:, :, None]    # (t, s, f)
    out = np.einsum("bsf,tsf->btf", xx.astype(np.complex128), g)
    out = out.reshape((xx.shape[0],) + dst.shape[:-1] + (xx.shape[2],))
    return out[0] if squeeze else out
