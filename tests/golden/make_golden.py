"""Writes the golden fixtures tests/golden/das_*.npz.

Inputs are small seeded problems (ffdas_amd.workloads.small_random plus two hand-made
geometries); expected outputs come from oracle/brute.py, the numpy FP64 restatement of the
reference's definition (the reference itself is CUDA-only and cannot run here, and it ships
no golden vectors of its own).  Also writes SHA256SUMS of the fixtures.
Run:  python tests/golden/make_golden.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle.brute as brute  # noqa: E402
from ffdas_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent


def save(name, w, scale=1.0):
    kw = dict(srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices, scale=scale)
    out = brute.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, **kw)
    d = dict(x=w.x, srcpos=w.srcpos, dstpos=w.dstpos, offsets=w.offsets, weights=w.weights,
             wavenum=np.float64(w.wavenum), scale=np.float64(scale), out=out)
    if w.srcdir is not None:
        d["srcdir"] = w.srcdir
    if w.sparse_indices is not None:
        d["sparse_indices"] = w.sparse_indices
    np.savez_compressed(OUT / f"das_{name}.npz", **d)


def main():
    save("real", W.small_random(seed=100, B=2, C=5, T=3, S=64, N=(4, 5)))
    save("dir", W.small_random(seed=101, B=1, C=6, T=2, S=80, N=(6,), srcdir=True))
    save("complex", W.small_random(seed=102, B=2, C=4, T=2, S=64, N=(3, 3), complex_iq=True))
    save("sparse", W.small_random(seed=103, B=2, C=4, T=6, K=3, S=72, N=(5, 2), sparse=True))
    save("scale", W.small_random(seed=104, B=1, C=4, T=3, S=64, N=(7,)), scale=1.04)
    # checksums of what was written (tests/test_oracle.py checks them)
    import hashlib
    lines = [f"{hashlib.sha256(f.read_bytes()).hexdigest()}  {f.name}"
             for f in sorted(OUT.glob("*.npz"))]
    (OUT / "SHA256SUMS").write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
