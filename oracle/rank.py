"""CPU oracle for ffdas `truncate_rank` (TEST INFRASTRUCTURE ONLY: imported by tests/ and
bench.py's CPU-baseline leg, never by the product).

Restates the definition of docs/api/truncate_rank.md (reconstruction from the singular
components [start, stop), counted from the largest) with numpy's SVD in FP64 -- the same
reference the reference's own tests use (bindings/python/tests/test_truncate_rank.py:39-44),
applied to x reshaped to (x.shape[0], prod(x.shape[1:])) in C order
(bindings/python/ffdas/truncate_rank.py:55-77).

Pinning (tests/test_truncate_rank.py): matrices built from known orthonormal factors and
distinct singular values, whose truncations are known exactly; keep-all returns x; the two
halves of a split add up to x; a rank-1 matrix loses everything past component 0; the
filtered matrix has exactly the kept singular values; and the reference's own test cases.
"""
from __future__ import annotations

import numpy as np


def svd_filter(x, start: int, stop: int | None = None) -> np.ndarray:
    """x with only the singular components [start, stop) kept (FP64 / complex128)."""
    x = np.asarray(x)
    m = x.shape[0]
    n = int(np.prod(x.shape[1:]))
    p = min(m, n)
    stop = p if stop is None else stop
    a = x.reshape(m, n).astype(np.complex128 if np.iscomplexobj(x) else np.float64)
    u, s, vh = np.linalg.svd(a, full_matrices=False)
    y = (u[:, start:stop] * s[start:stop]) @ vh[start:stop, :]
    return y.reshape(x.shape)


def svd_filter_timed(x, start: int, stop: int | None = None):
    """svd_filter in the input's precision (the CPU baseline), with its wall time [s]."""
    import time
    x = np.asarray(x)
    m = x.shape[0]
    n = int(np.prod(x.shape[1:]))
    t0 = time.perf_counter()
    a = x.reshape(m, n)
    u, s, vh = np.linalg.svd(a, full_matrices=False)
    stop = min(m, n) if stop is None else stop
    y = (u[:, start:stop] * s[start:stop]) @ vh[start:stop, :]
    return y.reshape(x.shape), time.perf_counter() - t0
