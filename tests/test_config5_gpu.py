"""Batched sparse compounding (config 5 at reduced size, GPU): the three probes -- 64- and
256-element linear arrays and the 32x32 matrix -- with focused transmits and a
k-nearest-foci index table, in one das_batched call with FP16 storage, against the FP64
oracle; the sparse adjoint; and a captured graph of the batched call."""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2, to_dev

pytestmark = pytest.mark.gpu

SMALL = {
    "linear64": dict(n_frames=5, n_tx=16, k=4, scale=0.2, n_samples=2048),
    "linear256": dict(n_frames=5, n_tx=16, k=4, scale=0.1, n_samples=2048),
    "matrix1024": dict(n_frames=4, n_tx=4, k=2, scale=0.1, n_samples=1536, n_tissue=10,
                       n_moving=6),
}


@pytest.fixture(scope="module")
def probes():
    return {name: W.config5(seed=50 + i, probe=name, **kw) for i, (name, kw) in
            enumerate(SMALL.items())}


def _problem(w, dev):
    d = to_dev(w, dev)
    return dict(x=d["x"], srcpos=d["srcpos"], dstpos=d["dstpos"], offsets=d["offsets"],
                weights=d["weights"], sparse_indices=d["sparse_indices"])


def test_config5_probes_in_one_batched_call(dev, probes):
    import torch
    import ffdas_amd
    ws = list(probes.values())
    outs = ffdas_amd.das_batched([_problem(w, dev) for w in ws], use_fp16=True)
    for w, o in zip(ws, outs):
        got = o.cpu().numpy()
        xq = torch.from_numpy(w.x).half().double().numpy()
        ref_q = oracle.das(xq, w.srcpos, w.dstpos, w.offsets, w.weights,
                           sparse_indices=w.sparse_indices)
        ref = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights,
                         sparse_indices=w.sparse_indices)
        assert rel_l2(got, ref_q) < 1e-4, w.name     # FP32 accumulation of FP16 samples
        assert rel_l2(got, ref) < 1e-3, w.name       # against the unquantised data


def test_config5_sparse_adjoint(dev, probes):
    import torch
    import ffdas_amd
    w = probes["linear64"]
    d = to_dev(w, dev)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((w.x.shape[0],) + w.dstpos.shape[:-1]).astype(np.float32)
    ref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, g,
                             sparse_indices=w.sparse_indices)
    gx, go, gw, gs = ffdas_amd.das_backward(d["x"], d["srcpos"], d["dstpos"], d["offsets"],
                                            d["weights"], torch.from_numpy(g).to(dev),
                                            sparse_indices=d["sparse_indices"])
    assert rel_l2(gx.cpu().numpy(), ref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), ref["grad_offsets"]) < 1e-4
    assert rel_l2(gw.cpu().numpy(), ref["grad_weights"]) < 1e-5
    assert abs(float(gs) - ref["grad_scale"]) <= 1e-4 * max(1.0, abs(ref["grad_scale"]))


def test_config5_graph_replay(dev, probes):
    import torch
    import ffdas_amd
    probs = [_problem(w, dev) for w in probes.values()]
    cap = ffdas_amd.CapturedCall(lambda: ffdas_amd.das_batched(probs, use_fp16=True))
    for k in range(3):
        for q in probs:
            q["x"].mul_(-0.8)
        got = [o.clone() for o in cap.replay()]
        torch.cuda.synchronize()
        eager = ffdas_amd.das_batched(probs, use_fp16=True)
        torch.cuda.synchronize()
        for a, b in zip(got, eager):
            assert torch.equal(a, b)
