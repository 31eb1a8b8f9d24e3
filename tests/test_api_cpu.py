"""Python operator checks that need no GPU: the reference's ValueError tests
(bindings/python/tests/test_das.py:39-77) and the absence of any CPU fallback."""
import pytest
import torch

import ffdas_amd


def _das_args(**overrides):
    n_ch, n_seq, n_samp, n_tgt = 4, 2, 16, 8
    args = dict(x=torch.zeros((n_ch, n_seq, n_samp)), srcpos=torch.zeros((n_ch, 3)),
                dstpos=torch.zeros((n_tgt, 3)), offsets=torch.zeros((n_seq, n_tgt)),
                weights=torch.ones((n_seq, n_tgt)))
    args.update(overrides)
    return args


def _das_sparse_args(**overrides):
    n_ch, n_seq, n_samp, n_tgt, n_sparse = 4, 8, 16, 6, 3
    args = dict(x=torch.zeros((n_ch, n_seq, n_samp)), srcpos=torch.zeros((n_ch, 3)),
                dstpos=torch.zeros((n_tgt, 3)), offsets=torch.zeros((n_sparse, n_tgt)),
                weights=torch.ones((n_sparse, n_tgt)),
                sparse_indices=torch.zeros((n_sparse, n_tgt), dtype=torch.int32))
    args.update(overrides)
    return args


def test_das_dstpos_wrong_trailing_dim():
    with pytest.raises(ValueError):
        ffdas_amd.das(**_das_args(dstpos=torch.zeros((8, 2))))


def test_das_dstpos_1d():
    with pytest.raises(ValueError):
        ffdas_amd.das(**_das_args(dstpos=torch.zeros((3,))))


def test_das_offsets_shape_mismatch():
    with pytest.raises(ValueError):
        ffdas_amd.das(**_das_args(offsets=torch.zeros((2, 99))))


def test_das_weights_shape_mismatch():
    with pytest.raises(ValueError):
        ffdas_amd.das(**_das_args(weights=torch.zeros((2, 99))))


def test_das_sparse_indices_shape_mismatch():
    with pytest.raises(ValueError):
        ffdas_amd.das_sparse(**_das_sparse_args(sparse_indices=torch.zeros((99, 6),
                                                                           dtype=torch.int32)))


def test_das_sparse_offsets_count_mismatch():
    with pytest.raises(ValueError):
        ffdas_amd.das_sparse(**_das_sparse_args(offsets=torch.zeros((5, 6))))


def test_host_tensors_are_rejected_not_computed():
    """There is no CPU implementation: host tensors fail loudly."""
    with pytest.raises(TypeError, match="ROCm device tensor"):
        ffdas_amd.das(**_das_args())


def test_view_checks():
    with pytest.raises(ValueError):
        ffdas_amd.view(torch.zeros((4, 3)), ffdas_amd.half2)
    with pytest.raises(ValueError):
        ffdas_amd.view(torch.zeros((4, 2), dtype=torch.float16), ffdas_amd.bfloat162)
    v = ffdas_amd.view(torch.zeros((4, 5, 2), dtype=torch.float16), ffdas_amd.half2)
    assert v.shape == (4, 5) and v.strides == (5, 1)
