"""The greens oracle (oracle/greens.py) pinned on the CPU, and the wrapper's shape errors.

Pins: the reference's test formula (bindings/python/tests/test_greens.py:8-16) recomputed
independently; one source gives one spherical wave x e^{ikr}/r (closed form); linearity in x;
and the far-field limit of a small source patch."""
import numpy as np
import pytest

import oracle.greens as G


def _reference_formula(srcpos, wavenums, x, dstpos):
    flat = dstpos.reshape(-1, 3)
    diff = srcpos[None, :, :] - flat[:, None, :]
    r = np.maximum(np.sqrt(np.sum(diff ** 2, axis=-1)), 1e-10)
    kern = np.exp(1j * wavenums[None, None, :] * r[:, :, None]) / r[:, :, None]
    out = np.einsum("bsf,tsf->btf", x, kern)
    return out.reshape(x.shape[0], *dstpos.shape[:-1], x.shape[2])


def test_matches_the_reference_test_formula():
    rng = np.random.default_rng(0)
    src = rng.standard_normal((4, 3)).astype("float32")
    dst = rng.standard_normal((6, 3)).astype("float32")
    k = rng.uniform(1, 10, 8).astype("float32")
    x = (rng.standard_normal((2, 4, 8)) + 1j * rng.standard_normal((2, 4, 8))).astype("complex64")
    # the reference's formula evaluated in FP64 (its test runs it on the FP32 arrays)
    want = _reference_formula(src.astype(np.float64), k.astype(np.float64),
                              x.astype(np.complex128), dst.astype(np.float64))
    np.testing.assert_allclose(G.greens(src, k, x, dst), want, rtol=1e-12)


def test_one_source_is_a_spherical_wave():
    src = np.array([[0.1, -0.2, 0.3]])
    dst = np.array([[1.0, 2.0, 2.3], [0.1, -0.2, 1.3]])
    k = np.array([2.0, -5.0])
    x = np.array([[[1.0 + 0j, 2.0 - 1j]]])
    r = np.array([np.sqrt(0.81 + 4.84 + 4.0), 1.0])
    want = x[0, 0][None, :] * np.exp(1j * k[None, :] * r[:, None]) / r[:, None]
    np.testing.assert_allclose(G.greens(src, k, x, dst)[0], want, rtol=1e-14)


def test_linear_in_x():
    rng = np.random.default_rng(1)
    src, dst = rng.standard_normal((5, 3)), rng.standard_normal((3, 3)) + 4
    k = rng.uniform(-5, 5, 4)
    a, b = rng.standard_normal((1, 5, 4)), 1j * rng.standard_normal((1, 5, 4))
    np.testing.assert_allclose(G.greens(src, k, 2 * a + 3 * b, dst),
                               2 * G.greens(src, k, a, dst) + 3 * G.greens(src, k, b, dst),
                               rtol=1e-12)


def test_far_field_of_a_small_patch():
    rng = np.random.default_rng(2)
    src = 1e-4 * rng.standard_normal((16, 3))
    dst = np.array([[0.0, 0.0, 50.0]])
    k = np.array([0.3])
    x = np.ones((1, 16, 1))
    got = G.greens(src, k, x, dst)[0, 0, 0]
    want = 16 * np.exp(1j * 0.3 * 50.0) / 50.0
    assert abs(got - want) / abs(want) < 1e-5


def test_wrapper_shape_errors():
    torch = pytest.importorskip("torch")
    import ffdas_amd
    with pytest.raises(ValueError):
        ffdas_amd.greens(torch.zeros((4, 3)), torch.ones(8), torch.zeros((1, 1, 4, 8),
                                                                          dtype=torch.complex64),
                         torch.zeros((6, 3)))
