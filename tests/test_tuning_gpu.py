"""Autotuner and tuned launch parameters (GPU): every configuration the tuner can pick must
give the oracle's result, and a tuned signature must be used by later calls."""
import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _problem(cplx=False, sparse=False, B=16):
    return W.small_random(seed=300 + 2 * cplx + sparse, B=B, C=12, T=4, K=3 if sparse else None,
                          S=300, N=(24, 40), complex_iq=cplx, sparse=sparse, srcdir=not cplx)


def _key_for(w, dev, tmp_path, **kw):
    """The tuning-cache key of this problem, read back from a saved cache."""
    import ffdas_amd
    ffdas_amd.clear_tuning()
    d = to_dev(w, dev)
    ffdas_amd.autotune(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"],
                       srcdir=d.get("srcdir"), wavenum=d.get("wavenum", 0.0),
                       sparse_indices=d.get("sparse_indices"), reps=1, **kw)
    f = tmp_path / "k.txt"
    ffdas_amd.save_tuning(f)
    lines = f.read_text().split("\n")
    assert len([l for l in lines if l]) == 1
    return lines[0].split()[0]


def _check(w, dev, **kw):
    import torch
    import ffdas_amd
    d = to_dev(w, dev)
    args = (d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"])
    opt = dict(srcdir=d.get("srcdir"), wavenum=d.get("wavenum", 0.0),
               sparse_indices=d.get("sparse_indices"))
    out = ffdas_amd.das(*args, srcdir=opt["srcdir"], wavenum=opt["wavenum"], **kw) \
        if opt["sparse_indices"] is None else \
        ffdas_amd.das_sparse(*args, opt["sparse_indices"], srcdir=opt["srcdir"],
                             wavenum=opt["wavenum"], **kw)
    ref = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, srcdir=w.srcdir,
                     wavenum=w.wavenum, sparse_indices=w.sparse_indices)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5
    rng = np.random.default_rng(4)
    g = rng.standard_normal(ref.shape)
    if np.iscomplexobj(ref):
        g = g + 1j * rng.standard_normal(ref.shape)
    g = g.astype(np.complex64 if np.iscomplexobj(ref) else np.float32)
    aref = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, g, srcdir=w.srcdir,
                              wavenum=w.wavenum, sparse_indices=w.sparse_indices)
    gx, go, gw, gs = ffdas_amd.das_backward(*args, torch.from_numpy(g).to(dev), **opt, **kw)
    assert rel_l2(gx.cpu().numpy(), aref["grad_x"]) < 1e-5
    assert rel_l2(go.cpu().numpy(), aref["grad_offsets"]) < 1e-4
    assert rel_l2(gw.cpu().numpy(), aref["grad_weights"]) < 1e-5
    assert abs(float(gs) - aref["grad_scale"]) <= 1e-4 * max(1.0, abs(aref["grad_scale"]))


def test_autotune_picks_and_uses_a_configuration(dev, tmp_path):
    import ffdas_amd
    w = _problem()
    ffdas_amd.clear_tuning()
    d = to_dev(w, dev)
    res = ffdas_amd.autotune(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"],
                             srcdir=d["srcdir"], reps=2)
    assert res["candidates"] > 10
    assert 0 < res["forward_ms"] <= res["default_forward_ms"]
    assert 0 < res["backward_ms"] <= res["default_backward_ms"]
    assert res["tile_h"] * res["tile_w"] in (0, 64)
    assert ffdas_amd.tuning_entries() == 1
    _check(w, dev)                      # runs with the tuned parameters
    ffdas_amd.clear_tuning()


CONFIGS = [
    # groups fi_frames tile_h tile_w param_groups bwdx_frames bwdx_lg bwdx_threads
    (1, 16, 8, 8, 1, 16, 4, 1024),
    (2, 8, 16, 4, 2, 8, 6, 512),
    (8, 4, 4, 16, 8, 4, 2, 1024),
    (4, 16, 32, 2, 4, 2, 5, 512),
    (8, 8, 2, 32, 0, 16, 0, 1024),
]


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("kind", ["real", "complex", "sparse"])
def test_forced_configurations_match_the_oracle(dev, tmp_path, cfg, kind):
    import ffdas_amd
    w = _problem(cplx=kind == "complex", sparse=kind == "sparse")
    key = _key_for(w, dev, tmp_path)
    f = tmp_path / "t.txt"
    f.write_text(key + " " + " ".join(str(v) for v in cfg) + "\n")
    ffdas_amd.clear_tuning()
    ffdas_amd.load_tuning(f)
    try:
        _check(w, dev)
    finally:
        ffdas_amd.clear_tuning()
