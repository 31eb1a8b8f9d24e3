"""Batched mixed-problem forward (ffdas_das_batched) and graph capture (GPU): every problem of
a batch must equal its own oracle result, and every replay of a captured graph must equal
the eager result of the same inputs."""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _mixed():
    return [
        W.small_random(seed=401, B=16, C=6, T=3, S=120, N=(10, 12)),
        W.small_random(seed=402, B=16, C=5, T=4, K=3, S=100, N=(3, 4, 5), sparse=True),
        W.small_random(seed=403, B=16, C=8, T=2, S=90, N=(20,), srcdir=True),
        W.small_random(seed=404, B=2, C=4, T=2, S=64, N=(6, 6)),        # natural layout
        W.small_random(seed=405, B=8, C=7, T=3, S=110, N=(9, 7)),       # another chunking
    ]


def _as_problem(w, dev, x_dtype=None):
    d = to_dev(w, dev, x_dtype)
    q = dict(x=d["x"], srcpos=d["srcpos"], dstpos=d["dstpos"], offsets=d["offsets"],
             weights=d["weights"])
    for k in ("srcdir", "sparse_indices", "wavenum"):
        if k in d:
            q[k] = d[k]
    return q


def _ref(w, x=None):
    return oracle.das(w.x if x is None else x, w.srcpos, w.dstpos, w.offsets, w.weights,
                      srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices)


def test_batched_mixed_problems_match_the_oracle(dev):
    import ffdas_amd
    ws = _mixed()
    outs = ffdas_amd.das_batched([_as_problem(w, dev) for w in ws])
    for w, o in zip(ws, outs):
        assert rel_l2(o.cpu().numpy(), _ref(w)) < 1e-5


def test_batched_fp16_and_complex(dev):
    import torch
    import ffdas_amd
    ws = [W.small_random(seed=410 + i, B=16, C=5, T=3, S=100, N=(8, 9 + i), complex_iq=True)
          for i in range(3)]
    outs = ffdas_amd.das_batched([_as_problem(w, dev) for w in ws], use_fp16=True)
    for w, o in zip(ws, outs):
        xq = torch.from_numpy(w.x).to(torch.complex32).to(torch.complex128).numpy()
        assert rel_l2(o.cpu().numpy(), _ref(w, xq)) < 1e-4


def test_batched_more_problems_than_one_launch_holds(dev):
    import ffdas_amd
    ws = [W.small_random(seed=420 + i, B=4, C=3 + i % 3, T=2, S=80, N=(5 + i, 6)) for i in range(14)]
    outs = ffdas_amd.das_batched([_as_problem(w, dev) for w in ws])
    for w, o in zip(ws, outs):
        assert rel_l2(o.cpu().numpy(), _ref(w)) < 1e-5


def test_batched_equals_single_calls_bitwise(dev, monkeypatch):
    """One launch for several problems computes each exactly as its own call does (with the
    same channel-group split, which fixes the summation order)."""
    import torch
    import ffdas_amd
    monkeypatch.setenv("FFDAS_CGROUPS", "4")
    ws = _mixed()[:3]
    probs = [_as_problem(w, dev) for w in ws]
    outs = ffdas_amd.das_batched(probs)
    for q, o in zip(probs, outs):
        if "sparse_indices" in q:
            single = ffdas_amd.das_sparse(q["x"], q["srcpos"], q["dstpos"], q["offsets"],
                                          q["weights"], q["sparse_indices"])
        else:
            single = ffdas_amd.das(q["x"], q["srcpos"], q["dstpos"], q["offsets"], q["weights"],
                                   srcdir=q.get("srcdir"))
        assert torch.equal(o, single)


def test_graph_replays_of_a_batched_call(dev):
    """Capture das_batched once, replay three times with the inputs changed in place and
    eager work in between; every replay equals an eager call on the same inputs."""
    import torch
    import ffdas_amd
    ws = _mixed()[:3]
    probs = [_as_problem(w, dev) for w in ws]
    cap = ffdas_amd.CapturedCall(lambda: ffdas_amd.das_batched(probs, use_fp16=True))
    for k in range(3):
        for q in probs:
            q["x"].mul_(1.5 if k % 2 == 0 else -0.5)
        outs = [o.clone() for o in cap.replay()]
        torch.cuda.synchronize()
        eager = ffdas_amd.das_batched(probs, use_fp16=True)     # eager work between replays
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e)


def test_graph_replays_of_forward_and_adjoint(dev):
    import torch
    import ffdas_amd
    w = W.small_random(seed=430, B=16, C=6, T=3, S=120, N=(12, 10))
    d = to_dev(w, dev)
    args = (d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    g = torch.randn((16, 12, 10), device=dev)

    def step():
        out = ffdas_amd.das(*args)
        gx, go, _, gs = ffdas_amd.das_backward(*args, g, need_weights=False)
        return out, gx, go, gs

    cap = ffdas_amd.CapturedCall(step)
    for k in range(3):
        g.mul_(0.7)
        d["x"].add_(0.1)
        got = [t.clone() for t in cap.replay()]
        torch.cuda.synchronize()
        ref = step()
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
