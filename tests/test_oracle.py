"""Pins the CPU oracle (oracle/das_oracle.cpp, oracle/brute.py) without restating it.

Every check below is chosen so that a plausible mistake in the oracle (an off-by-one sample
index, a wrong row/transmit index, a dropped clamp, a sign error in the phase, a transposed
offsets layout, a non-strict directivity test, a wrong sparse lookup, a wrong adjoint) fails
at least one of them:
  * the reference's known answers (bindings/python/tests/test_das.py:83-132);
  * closed forms on geometries whose distances are exact by construction (on-axis voxels,
    3-4-5 triangles), linear interpolation against numpy.interp;
  * invariants (linearity, additivity over transmits, batch independence, sparse == dense,
    top-K sparse == dense (bindings/python/examples/sparse_compounding.py:131-159),
    channels_trailing, beta accumulation, delay-scale equivalence);
  * the separately written numpy brute force and the committed golden fixtures;
  * adjoint: <A x, y> = <x, A^T y>, and central differences of the forward.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
import oracle.brute as brute
from ffdas_amd import workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden"
R0 = math.sqrt(float(np.float32(1e-7)))   # clamped radius at zero distance


def _single(x_row, tau, w=1.0, S=None, wavenum=0.0, srcdir=None, dst=(0.0, 0.0, 0.0)):
    """One channel at the origin, one transmit, one voxel; offset chosen so that
    tau = offset + r equals `tau` with r = |dst| (exact for the geometries used)."""
    x_row = np.asarray(x_row)
    S = S or x_row.shape[-1]
    r = float(np.linalg.norm(dst)) if any(dst) else R0
    x = x_row.reshape(1, 1, S)
    srcpos = np.zeros((1, 3), np.float32)
    dstpos = np.asarray(dst, np.float32).reshape(1, 3)
    off = np.array([[tau - r]], np.float64)
    assert float(np.float32(off[0, 0])) == off[0, 0], "offset must be exact in float32"
    return oracle.das(x, srcpos, dstpos, off.astype(np.float32), np.full((1, 1), w, np.float32),
                      wavenum=wavenum, srcdir=srcdir)[0]


# ---- reference known answers ------------------------------------------------------------

def test_impulse_known_answer():
    """test_das.py:121-132: x[32] = 1, everything at the origin, offset 32 -> 1 - sqrt(1e-7)."""
    x = np.zeros((1, 1, 64), np.float32)
    x[0, 0, 32] = 1
    out = oracle.das(x, np.zeros((1, 3), np.float32), np.zeros((1, 3), np.float32),
                     np.full((1, 1), 32.0, np.float32), np.ones((1, 1), np.float32))
    assert abs(out[0] - (1 - R0)) < 1e-13
    assert abs(out[0]) > 0.5


def test_zero_input_and_zero_weights():
    w = W.small_random(seed=3)
    z = oracle.das(np.zeros_like(w.x), w.srcpos, w.dstpos, w.offsets, w.weights)
    assert np.all(z == 0)
    z = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, np.zeros_like(w.weights))
    assert np.all(z == 0)


def test_output_shapes():
    x3 = np.zeros((4, 2, 32), np.float32)
    o = oracle.das(x3, np.zeros((4, 3), np.float32), np.zeros((8, 6, 3), np.float32),
                   np.zeros((2, 8, 6), np.float32), np.ones((2, 8, 6), np.float32))
    assert o.shape == (8, 6)
    x4 = np.zeros((3, 4, 2, 32), np.float32)
    o = oracle.das(x4, np.zeros((4, 3), np.float32), np.zeros((8, 3), np.float32),
                   np.zeros((2, 8), np.float32), np.ones((2, 8), np.float32))
    assert o.shape == (3, 8)


# ---- closed forms -----------------------------------------------------------------------

@pytest.mark.parametrize("tau", [0.0, 0.25, 7.5, 31.875, 62.0, 62.96875])
def test_two_tap_interpolation_matches_numpy_interp(tau):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(64)
    got = _single(x, tau, dst=(3.0, 4.0, 0.0))
    assert abs(got - np.interp(tau, np.arange(64), x)) < 1e-12


@pytest.mark.parametrize("tau", [-0.5, -2.0 ** -10, 63.0, 63.5, 100.0])
def test_out_of_range_is_zero(tau):
    """0 <= floor(tau) < S-1, else zero (das_alg1_kernels.cuh:70)."""
    x = np.ones(64)
    assert _single(x, tau, dst=(0.0, 0.0, 5.0)) == 0.0


def test_ramp_closed_form():
    """x[k] = a k + c on every row -> out = sum_in w (a tau + c), with on-axis voxels whose
    distances are exact (voxel (0,0,z) to a channel at the origin)."""
    S, a, c = 128, 0.37, -1.25
    z = np.array([2.5, 10.0, 33.25, 60.0], np.float32)
    C, T = 1, 3
    x = np.broadcast_to(a * np.arange(S) + c, (C, T, S)).copy()
    dst = np.stack([np.zeros(4), np.zeros(4), z], -1).astype(np.float32)
    off = np.array([[0.5, 1.0, 2.25, 80.0], [10.0, -2.0, 0.0, 3.5], [5.125, 5.125, 5.125, 5.125]],
                   np.float32)
    wts = np.array([[1.0, 0.5, 2.0, 1.0], [0.25, 1.0, 1.0, 1.0], [3.0, 1.0, 0.5, 2.0]], np.float32)
    out = oracle.das(x, np.zeros((1, 3), np.float32), dst, off, wts)
    tau = off.astype(np.float64) + z.astype(np.float64)[None, :]
    inr = (np.floor(tau) >= 0) & (np.floor(tau) < S - 1)
    expect = np.sum(np.where(inr, wts * (a * tau + c), 0.0), axis=0)
    assert np.allclose(out, expect, rtol=0, atol=1e-12)


def test_constant_complex_closed_form():
    """Complex constant c: out = c * sum_in w exp(i k tau) (das_alg1_kernels.cuh:81-85)."""
    S, kap, c = 64, 0.7, 0.3 - 0.8j
    x = np.full((1, 2, S), c, np.complex128)
    dst = np.array([[3.0, 4.0, 0.0]], np.float32)               # r = 5 exactly
    off = np.array([[1.25], [20.5]], np.float32)
    wts = np.array([[0.5], [2.0]], np.float32)
    out = oracle.das(x, np.zeros((1, 3), np.float32), dst, off, wts, wavenum=kap)
    tau = off[:, 0].astype(np.float64) + 5.0
    expect = c * np.sum(wts[:, 0] * np.exp(1j * np.float64(np.float32(kap)) * tau))
    assert abs(out[0] - expect) < 1e-12


def test_directivity_is_strict():
    """proj > srcdir.w strictly (das_alg1_kernels.cuh:75): proj == 0 with w == 0 excluded."""
    x = np.ones((1, 1, 64))
    sd = np.array([[0.0, 0.0, 1.0, 0.0]], np.float32)
    assert _single(x, 10.0, dst=(5.0, 0.0, 0.0), srcdir=sd) == 0.0        # proj == 0
    assert _single(x, 10.0, dst=(4.0, 0.0, 3.0), srcdir=sd) == 1.0        # proj == 0.6
    sd2 = np.array([[0.0, 0.0, 1.0, -1.0]], np.float32)
    assert _single(x, 10.0, dst=(0.0, 0.0, -5.0), srcdir=sd2) == 0.0      # proj == -1


# ---- invariants -------------------------------------------------------------------------

def test_linearity_in_x_and_weights():
    w = W.small_random(seed=4, srcdir=True)
    rng = np.random.default_rng(5)
    y = rng.standard_normal(w.x.shape)
    f = lambda xx, ww=w.weights: oracle.das(xx, w.srcpos, w.dstpos, w.offsets, ww,
                                            srcdir=w.srcdir)
    assert np.allclose(f(2.0 * w.x - 0.5 * y), 2.0 * f(w.x) - 0.5 * f(y), atol=1e-11)
    w2 = rng.uniform(-1, 1, w.weights.shape).astype(np.float32)
    assert np.allclose(f(w.x, w.weights + w2), f(w.x) + f(w.x, w2), atol=1e-11)


def test_additivity_over_transmits():
    """Compounding = sum of single-transmit images (examples/compound.py:117-140)."""
    w = W.small_random(seed=6, T=4)
    full = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    parts = sum(oracle.das(w.x[:, :, i:i + 1], w.srcpos, w.dstpos, w.offsets[i:i + 1],
                           w.weights[i:i + 1]) for i in range(4))
    assert np.allclose(full, parts, atol=1e-11)


def test_batch_independence():
    w = W.small_random(seed=7, B=3)
    full = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    for b in range(3):
        one = oracle.das(w.x[b], w.srcpos, w.dstpos, w.offsets, w.weights)
        assert np.array_equal(full[b], one)


def test_sparse_all_transmits_equals_dense():
    w = W.small_random(seed=8, T=4)
    N = w.dstpos.shape[:-1]
    idx = np.broadcast_to(np.arange(4).reshape((4,) + (1,) * len(N)), (4,) + N).astype(np.int32)
    dense = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    sp = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, sparse_indices=idx)
    assert np.allclose(dense, sp, atol=1e-12)
    # permuted slots give the same sum
    perm = idx[::-1].copy()
    sp2 = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets[::-1].copy(), w.weights[::-1].copy(),
                     sparse_indices=perm)
    assert np.allclose(dense, sp2, atol=1e-11)


def test_topk_sparse_equals_dense_when_rest_is_zero():
    """examples/sparse_compounding.py:131-159: zero-weight transmits may be skipped."""
    w = W.small_random(seed=9, T=6)
    rng = np.random.default_rng(9)
    wt = w.weights.copy()
    keep = np.argsort(rng.uniform(size=(6,) + wt.shape[1:]), axis=0)[:2]     # 2 per voxel
    mask = np.zeros_like(wt, bool)
    np.put_along_axis(mask, keep, True, axis=0)
    wt[~mask] = 0
    dense = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, wt)
    idx = np.argpartition(np.abs(wt), -2, axis=0)[-2:].astype(np.int32)
    sp = oracle.das(w.x, w.srcpos, w.dstpos, np.take_along_axis(w.offsets, idx, 0),
                    np.take_along_axis(wt, idx, 0), sparse_indices=idx)
    assert np.allclose(dense, sp, atol=1e-11)


def test_out_of_range_sparse_index_contributes_nothing():
    w = W.small_random(seed=10, T=3, K=2, sparse=True)
    si = w.sparse_indices.copy()
    base = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, sparse_indices=si)
    si2 = si.copy()
    si2[1] = 7                                   # second slot out of range everywhere
    only0 = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets[:1], w.weights[:1],
                       sparse_indices=si[:1])
    got = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, sparse_indices=si2)
    assert np.allclose(got, only0, atol=1e-12)
    assert not np.allclose(base, only0)


def test_channels_trailing_equivalence():
    w = W.small_random(seed=11)
    a = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    b = oracle.das(np.swapaxes(w.x, 1, 2).copy(), w.srcpos, w.dstpos, w.offsets, w.weights,
                   channels_trailing=True)
    assert np.array_equal(a, b)


def test_beta_accumulation():
    w = W.small_random(seed=12)
    a = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    prev = np.random.default_rng(1).standard_normal(a.shape)
    b = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, beta=0.5, out=prev)
    assert np.allclose(b, 0.5 * prev + a, atol=1e-12)


def test_scale_equals_scaled_geometry():
    """tau = s (ofs + r): scaling every position and offset by s instead gives the same image."""
    w = W.small_random(seed=13)
    s = 2.0                                      # exact scaling in float32
    a = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, scale=s)
    b = oracle.das(w.x, (w.srcpos * s).astype(np.float32), (w.dstpos * s).astype(np.float32),
                   (w.offsets * s).astype(np.float32), w.weights)
    assert np.allclose(a, b, atol=1e-10)


# ---- brute force and golden fixtures ----------------------------------------------------

@pytest.mark.parametrize("kind", ["real", "complex", "sparse", "dir", "scale"])
def test_cpp_oracle_matches_numpy_brute_force(kind):
    kw = {}
    if kind == "real":
        w = W.small_random(seed=20)
    elif kind == "complex":
        w = W.small_random(seed=21, complex_iq=True)
    elif kind == "sparse":
        w = W.small_random(seed=22, T=5, K=3, sparse=True)
    elif kind == "dir":
        w = W.small_random(seed=23, srcdir=True)
    else:
        w = W.small_random(seed=24)
        kw["scale"] = 0.97
    args = (w.x, w.srcpos, w.dstpos, w.offsets, w.weights)
    kw.update(srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices)
    assert np.allclose(oracle.das(*args, **kw), brute.das(*args, **kw), rtol=0, atol=1e-11)


def test_golden_fixtures():
    files = sorted(GOLDEN.glob("das_*.npz"))
    assert files, "golden fixtures missing (tests/golden/make_golden.py)"
    for f in files:
        z = np.load(f, allow_pickle=False)
        kw = dict(wavenum=float(z["wavenum"]), scale=float(z["scale"]))
        if "srcdir" in z:
            kw["srcdir"] = z["srcdir"]
        if "sparse_indices" in z:
            kw["sparse_indices"] = z["sparse_indices"]
        got = oracle.das(z["x"], z["srcpos"], z["dstpos"], z["offsets"], z["weights"], **kw)
        assert np.allclose(got, z["out"], rtol=0, atol=1e-10), f.name


# ---- adjoint ----------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["real", "complex", "sparse", "dir"])
def test_adjoint_dot_product(kind):
    w = {"real": W.small_random(seed=30), "complex": W.small_random(seed=31, complex_iq=True),
         "sparse": W.small_random(seed=32, T=5, K=3, sparse=True),
         "dir": W.small_random(seed=33, srcdir=True)}[kind]
    rng = np.random.default_rng(3)
    kw = dict(srcdir=w.srcdir, wavenum=w.wavenum, sparse_indices=w.sparse_indices, scale=1.1)
    ax = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, **kw)
    y = rng.standard_normal(ax.shape)
    if np.iscomplexobj(ax):
        y = y + 1j * rng.standard_normal(ax.shape)
    g = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, y, **kw)
    lhs = np.vdot(y, ax)
    rhs = np.vdot(g["grad_x"], w.x.astype(np.complex128 if np.iscomplexobj(w.x) else np.float64))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + 1)
    # and the numpy brute-force transpose agrees
    gx = brute.das_adjoint_x(w.x.shape, w.srcpos, w.dstpos, w.offsets, w.weights, y,
                             is_complex=np.iscomplexobj(w.x), **kw)
    assert np.allclose(g["grad_x"], gx, atol=1e-11)


@pytest.mark.parametrize("complex_iq", [False, True])
def test_adjoint_parameters_by_central_differences(complex_iq):
    """d/d offsets, d/d weights, d/d scale of L = Re<g, das(...)> by central differences of
    the (independently pinned) forward; h small enough not to cross an interpolation knot."""
    w = W.small_random(seed=40, B=2, C=3, T=2, N=(3, 2), complex_iq=complex_iq)
    rng = np.random.default_rng(4)
    shape = (2,) + w.dstpos.shape[:-1]
    gout = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if complex_iq else 0)
    s0 = 1.02

    def loss(off=w.offsets, wt=w.weights, s=s0):
        o = brute.das(w.x, w.srcpos, w.dstpos, off, wt, wavenum=w.wavenum, scale=s)
        return float(np.real(np.vdot(gout, o)))

    g = oracle.das_adjoint(w.x, w.srcpos, w.dstpos, w.offsets.astype(np.float64),
                           w.weights, gout, wavenum=w.wavenum, scale=s0)
    h = 1e-6
    off64 = w.offsets.astype(np.float64)
    for idx in [(0, 0, 0), (1, 2, 1), (0, 1, 1)]:
        e = np.zeros_like(off64)
        e[idx] = h
        fd = (loss(off=off64 + e) - loss(off=off64 - e)) / (2 * h)
        assert abs(fd - g["grad_offsets"][idx]) < 1e-5 * max(1, abs(fd)), idx
        wt64 = w.weights.astype(np.float64)
        fdw = (loss(wt=wt64 + e) - loss(wt=wt64 - e)) / (2 * h)
        assert abs(fdw - g["grad_weights"][idx]) < 1e-6 * max(1, abs(fdw)), idx
    fds = (loss(s=s0 + h) - loss(s=s0 - h)) / (2 * h)
    assert abs(fds - g["grad_scale"]) < 1e-4 * max(1, abs(fds))


# ---- the FP32 build of the oracle (CPU baseline) --------------------------------------

def test_fp32_oracle_error_is_bounded():
    """The plain FP32 restatement (the CPU baseline) stays within 5e-5 of FP64 on config 1."""
    w = W.config1()
    a = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, bits=64)
    b = oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights, bits=32)
    assert oracle.rel_l2(b, a) < 5e-5


def test_golden_fixture_checksums():
    """The committed fixtures are the ones tests/golden/make_golden.py wrote (SHA256SUMS)."""
    import hashlib
    from pathlib import Path
    gold = Path(__file__).resolve().parent / "golden"
    lines = [l.split() for l in (gold / "SHA256SUMS").read_text().splitlines() if l.strip()]
    assert len(lines) == len(list(gold.glob("*.npz")))
    for digest, name in lines:
        assert hashlib.sha256((gold / name).read_bytes()).hexdigest() == digest, name
