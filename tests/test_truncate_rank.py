"""truncate_rank on the CPU: the oracle pinned against matrices with known singular factors
and the reference's own cases, and the Python layer's argument errors
(bindings/python/tests/test_truncate_rank.py:10-34)."""
import numpy as np
import pytest
import torch

import oracle.rank as R


def _factors(m, n, cplx, rng, s=None):
    """x = U diag(s) V^H with orthonormal U (m x p), V (n x p) and distinct s."""
    p = min(m, n)

    def orth(k):
        a = rng.standard_normal((k, p))
        if cplx:
            a = a + 1j * rng.standard_normal((k, p))
        q, _ = np.linalg.qr(a)
        return q

    u, v = orth(m), orth(n)
    s = np.sort(rng.uniform(1.0, 10.0, p))[::-1] if s is None else np.asarray(s, float)
    return u, s, v, (u * s) @ v.conj().T


@pytest.mark.parametrize("m, n", [(6, 20), (20, 6), (9, 9)])
@pytest.mark.parametrize("cplx", [False, True])
def test_oracle_known_factors(m, n, cplx):
    rng = np.random.default_rng(m * 31 + n + cplx)
    u, s, v, x = _factors(m, n, cplx, rng)
    p = min(m, n)
    for a, b in [(0, 1), (1, p), (2, p - 1), (0, p)]:
        want = (u[:, a:b] * s[a:b]) @ v[:, a:b].conj().T
        assert np.allclose(R.svd_filter(x, a, b), want, atol=1e-10)


def test_oracle_keep_all_split_and_rank_one():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((8, 4, 6))
    assert np.allclose(R.svd_filter(x, 0), x, atol=1e-12)
    assert np.allclose(R.svd_filter(x, 0, 3) + R.svd_filter(x, 3), x, atol=1e-12)
    a = np.outer(rng.standard_normal(5), rng.standard_normal(7))
    assert np.allclose(R.svd_filter(a, 1), 0.0, atol=1e-12)
    assert np.allclose(R.svd_filter(a, 0, 1), a, atol=1e-12)


def test_oracle_keeps_exactly_the_selected_singular_values():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((12, 30)) + 1j * rng.standard_normal((12, 30))
    s = np.linalg.svd(x, compute_uv=False)
    y = R.svd_filter(x, 3, 9)
    sy = np.linalg.svd(y, compute_uv=False)
    assert np.allclose(sy[:6], s[3:9], atol=1e-10)
    assert np.allclose(sy[6:], 0.0, atol=1e-10)


def test_oracle_reference_shape_cases():
    rng = np.random.default_rng(3)
    assert R.svd_filter(rng.standard_normal((8, 4, 16)), 0, 4).shape == (8, 4, 16)
    x = rng.standard_normal((64, 8))
    assert np.allclose(R.svd_filter(x, 0, 8), x, atol=1e-12)


@pytest.mark.parametrize("shape, start, stop", [
    ((10,), 0, None),          # 1-D rejected
    ((4, 8), -1, None),        # start negative
    ((4, 8), 4, None),         # start too large
    ((4, 8), 2, 1),            # stop before start
    ((4, 8), 0, 5),            # stop too large
])
def test_argument_errors(shape, start, stop):
    import ffdas_amd
    with pytest.raises(ValueError):
        ffdas_amd.truncate_rank(torch.ones(shape), start=start, stop=stop)


def test_out_shape_error():
    import ffdas_amd
    with pytest.raises(ValueError):
        ffdas_amd.truncate_rank(torch.ones((4, 8)), 0, out=torch.empty((8, 4)))
