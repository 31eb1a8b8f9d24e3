"""Forward accuracy per storage type against the FP64 oracle (configurations 1 and 3: the
128-element array, three plane waves, 2048 samples, 256 x 128 voxels; 1 and 8 frames):
rel-L2 and max |error| / max |reference|, and, for the 16-bit storages, the same against
the oracle run on the storage-rounded data (the part of the error that is the kernel's own
FP32 arithmetic).  Printed as one table; the bounds are the stated targets (1e-5 FP32 and
for every 16-bit storage against its rounded data; 1e-3 FP16/half2 and 1.5e-3 BF16 -- its
8-bit mantissa's own rounding floor -- against the unrounded data)."""
import json

import numpy as np
import pytest

import oracle
from ffdas_amd import workloads as W
from tests.helpers import to_dev

pytestmark = pytest.mark.gpu


def _ref(w, x=None):
    return oracle.das(w.x if x is None else x, w.srcpos, w.dstpos, w.offsets, w.weights,
                      srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices)


def _err(got, want):
    got = np.asarray(got, dtype=np.complex128 if np.iscomplexobj(got) else np.float64)
    d = np.abs(got - want)
    return {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(want)),
            "max_abs_rel": float(d.max() / np.abs(want).max())}


def test_accuracy_table(dev):
    import torch
    import ffdas_amd

    def run(w, x_dtype=None, **kw):
        d = to_dev(w, dev, x_dtype)
        out = ffdas_amd.das(d["x"], d["srcpos"], d["dstpos"], d["offsets"], d["weights"],
                            srcdir=d.get("srcdir"), wavenum=d.get("wavenum", 0.0), **kw)
        return out.cpu().numpy()

    rows = {}
    w1 = W.config1()
    rows["config1 fp32"] = _err(run(w1), _ref(w1))
    rows["config1 fp64"] = _err(run(w1, torch.float64), _ref(w1, w1.x.astype(np.float64)))
    w3 = W.config3()
    r3 = _ref(w3)
    rows["config3 fp32"] = _err(run(w3), r3)
    q16 = w3.x.astype(np.float16).astype(np.float64)
    g16 = run(w3, use_fp16=True)
    rows["config3 fp16"] = dict(_err(g16, r3), vs_rounded=_err(g16, _ref(w3, q16)))
    qbf = torch.from_numpy(w3.x).to(torch.bfloat16).to(torch.float64).numpy()
    gbf = run(w3, use_bf16=True)
    rows["config3 bf16"] = dict(_err(gbf, r3), vs_rounded=_err(gbf, _ref(w3, qbf)))
    wc = W.config3(complex_iq=True)
    rc = _ref(wc)
    rows["config3 IQ complex64"] = _err(run(wc), rc)
    qc = torch.from_numpy(wc.x).to(torch.complex32).to(torch.complex128).numpy()
    gc = run(wc, torch.complex32)
    rows["config3 IQ complex32 (half2)"] = dict(_err(gc, rc), vs_rounded=_err(gc, _ref(wc, qc)))
    print("\naccuracy " + json.dumps(rows))
    for k, v in rows.items():
        bound = 1.5e-3 if "bf16" in k else (1e-3 if ("fp16" in k or "half2" in k) else 1e-5)
        assert v["rel_l2"] < bound, (k, v)
        if "vs_rounded" in v:
            assert v["vs_rounded"]["rel_l2"] < 1e-5, (k, v)
