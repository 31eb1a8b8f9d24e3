"""Seeded synthetic ultrasound workloads for the delay-and-sum path.

Every choice the workload definitions leave open is fixed here and recorded in the
returned ``meta`` dict (seed, pulse width, noise convention, grid extents, geometry
conventions), so that the CPU oracle and the GPU library see identical inputs.

Units follow the reference convention (docs/conventions.md:9-35): positions in sampling
wavelengths (c / fs), offsets in samples.  Plane-wave transmit delay is
(z*cos(theta) + x*sin(theta)) * fs/c (bindings/python/examples/compound.py:104-110);
diverging-wave delay is (|v - y| - |v - ref|) * fs/c (docs/getting-started/quickstart.md:44-53).

Echo model: each point scatterer p with amplitude A contributes to channel m, transmit i
  A * exp(-(k - tau)^2 / (2 sigma^2)) * cos(2 pi f0 (k - tau) / fs)
at sample k, with round-trip delay tau = tx_i(p) + |p - e_m| * fs/c, plus i.i.d. Gaussian
noise of standard deviation 0.01.  sigma = 4.8 samples (about two cycles at 5 MHz / 40 MHz).
Echoes are evaluated inside +-6 sigma windows only (the Gaussian is below 1.5e-8 outside).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_SOUND = 1540.0          # m/s
FS = 40e6                 # Hz
F0 = 5e6                  # Hz, pulse centre frequency
KS = FS / C_SOUND         # sampling wavelengths per metre (~25.97 per mm)
SIGMA = 4.8               # samples
NOISE_STD = 0.01
HALF_WIN = int(math.ceil(6 * SIGMA))


@dataclass
class Workload:
    """Inputs of one das / das_sparse call in the reference's Python layout."""
    name: str
    x: np.ndarray              # (B, C, T, S) float32 or complex64
    srcpos: np.ndarray         # (C, 3) float32, sampling wavelengths
    dstpos: np.ndarray         # (*spatial, 3) float32, sampling wavelengths
    offsets: np.ndarray        # (K, *spatial) float32, samples
    weights: np.ndarray        # (K, *spatial) float32
    srcdir: np.ndarray | None = None
    wavenum: float = 0.0
    sparse_indices: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def units(self) -> int:
        """Work units U = voxels * channels * compounding slots * frames."""
        B, C = self.x.shape[0], self.x.shape[1]
        K = self.offsets.shape[0]
        N = int(np.prod(self.dstpos.shape[:-1]))
        return N * C * K * B


def linear_array(n: int, pitch_m: float) -> np.ndarray:
    xs = (np.arange(n) - (n - 1) / 2.0) * pitch_m
    return np.stack([xs, np.zeros(n), np.zeros(n)], axis=-1)


def matrix_array(n: int, pitch_m: float) -> np.ndarray:
    e = (np.arange(n) - (n - 1) / 2.0) * pitch_m
    ex, ey = np.meshgrid(e, e, indexing="ij")
    return np.stack([ex.ravel(), ey.ravel(), np.zeros(n * n)], axis=-1)


def plane_wave_delay(points_m: np.ndarray, theta: np.ndarray) -> np.ndarray:
    """One-way plane-wave delay in samples, (T, *points)."""
    th = np.asarray(theta, np.float64).reshape((-1,) + (1,) * (points_m.ndim - 1))
    return (points_m[..., 2] * np.cos(th) + points_m[..., 0] * np.sin(th)) * KS


def diverging_delay(points_m: np.ndarray, sources_m: np.ndarray, ref_m: np.ndarray) -> np.ndarray:
    """(|v - y| - |v - ref|) * fs/c in samples, (T, *points)."""
    sh = (-1,) + (1,) * (points_m.ndim - 1) + (3,)
    v = sources_m.reshape(sh)
    d = np.linalg.norm(points_m[None] - v, axis=-1)
    return (d - ref_m.reshape((-1,) + (1,) * (points_m.ndim - 1))) * KS


def synth_echoes(rng, n_frames, elem_m, tx_delay_fn, n_tx, n_samples, scat_fn,
                 complex_iq=False):
    """Channel data (B, C, T, S).  scat_fn(b) -> (P,3) positions [m], (P,) amplitudes.
    tx_delay_fn(points (P,3)) -> (T, P) transmit delays in samples."""
    C = elem_m.shape[0]
    x = np.zeros((n_frames, C, n_tx, n_samples), dtype=np.float64)
    offs = np.arange(-HALF_WIN, HALF_WIN + 1)
    for b in range(n_frames):
        pos, amp = scat_fn(b)
        tx = tx_delay_fn(pos)                                            # (T, P)
        rx = np.linalg.norm(pos[None, :, :] - elem_m[:, None, :], axis=-1) * KS   # (C, P)
        tau = tx[None, :, :] + rx[:, None, :]                            # (C, T, P)
        k0 = np.floor(tau).astype(np.int64)
        kk = k0[..., None] + offs                                        # (C, T, P, W)
        dt = kk - tau[..., None]
        val = amp[None, None, :, None] * np.exp(-dt * dt / (2 * SIGMA * SIGMA)) * \
            np.cos(2 * np.pi * F0 * dt / FS)
        ok = (kk >= 0) & (kk < n_samples)
        ci = np.broadcast_to(np.arange(C)[:, None, None, None], kk.shape)
        ti = np.broadcast_to(np.arange(n_tx)[None, :, None, None], kk.shape)
        np.add.at(x[b], (ci[ok], ti[ok], kk[ok]), val[ok])
    x += rng.normal(0.0, NOISE_STD, size=x.shape)
    if complex_iq:
        # demodulate at F0: x_iq[k] = x[k] * exp(-2 pi i F0 k / fs) (phase reference at sample 0),
        # matching the wavenum = 2 pi F0 / fs rotation convention used by das.
        k = np.arange(n_samples)
        return (x * np.exp(-2j * np.pi * F0 * k / FS)).astype(np.complex64)
    return x.astype(np.float32)


def _grid2d(nz, nx, z_range_m=(5e-3, 45e-3), x_half_m=19.2e-3):
    z = np.linspace(z_range_m[0], z_range_m[1], nz)
    xs = np.linspace(-x_half_m, x_half_m, nx)
    zz, xx = np.meshgrid(z, xs, indexing="ij")
    return np.stack([xx, np.zeros_like(xx), zz], axis=-1)   # (nz, nx, 3) metres


def config1(seed: int = 1, n_frames: int = 1, complex_iq: bool = False,
            n_samples: int = 2048) -> Workload:
    """Plain 2-D das: 128-element linear array, pitch 0.3 mm, 3 plane waves (-10,0,+10 deg),
    2048 samples, 20 scatterers, grid 256(z) x 128(x) x 1, weights 1/3."""
    rng = np.random.default_rng(seed)
    elem = linear_array(128, 0.3e-3)
    thetas = np.deg2rad([-10.0, 0.0, 10.0])
    P = 20
    sx = rng.uniform(-15e-3, 15e-3, P)
    sz = rng.uniform(5e-3, 45e-3, P)
    amp = rng.uniform(0.5, 1.0, P)
    pos = np.stack([sx, np.zeros(P), sz], axis=-1)
    x = synth_echoes(rng, n_frames, elem, lambda p: plane_wave_delay(p, thetas), 3, n_samples,
                     lambda b: (pos, amp), complex_iq=complex_iq)
    grid = _grid2d(256, 128)[:, :, None, :]                   # (256, 128, 1, 3)
    offsets = plane_wave_delay(grid, thetas).astype(np.float32)
    weights = np.full(offsets.shape, 1.0 / 3.0, dtype=np.float32)
    wavenum = 2 * np.pi * F0 / FS if complex_iq else 0.0
    return Workload("config1_2d_das", x, (elem * KS).astype(np.float32),
                    (grid * KS).astype(np.float32), offsets, weights, wavenum=float(wavenum),
                    meta=dict(seed=seed, sigma=SIGMA, noise_std=NOISE_STD, f0=F0, fs=FS,
                              c=C_SOUND, grid="z 5..45 mm x +-19.2 mm", frames=n_frames))


def config3(seed: int = 3, n_frames: int = 8, complex_iq: bool = False) -> Workload:
    """Precision form of config 1 with 8 frames (independent noise)."""
    w = config1(seed=seed, n_frames=n_frames, complex_iq=complex_iq)
    w.name = "config3_precision_iq" if complex_iq else "config3_precision"
    return w


def config4(seed: int = 4, n_frames: int = 16, nz: int = 256, nx: int = 256,
            n_angles: int = 11) -> Workload:
    """Differentiation config: 128-element linear array, 11 plane waves (-15..+15 deg),
    2048 samples, 16 frames, grid 256 x 256 (z 5..45 mm, x +-19.2 mm), weights 1/11."""
    rng = np.random.default_rng(seed)
    elem = linear_array(128, 0.3e-3)
    thetas = np.deg2rad(np.linspace(-15.0, 15.0, n_angles))
    P = 20
    pos = np.stack([rng.uniform(-15e-3, 15e-3, P), np.zeros(P), rng.uniform(5e-3, 45e-3, P)], -1)
    amp = rng.uniform(0.5, 1.0, P)
    x = synth_echoes(rng, n_frames, elem, lambda p: plane_wave_delay(p, thetas), n_angles, 2048,
                     lambda b: (pos, amp))
    grid = _grid2d(nz, nx)
    offsets = plane_wave_delay(grid, thetas).astype(np.float32)
    weights = np.full(offsets.shape, 1.0 / n_angles, dtype=np.float32)
    return Workload("config4_fwd_adjoint", x, (elem * KS).astype(np.float32),
                    (grid * KS).astype(np.float32), offsets, weights,
                    meta=dict(seed=seed, sigma=SIGMA, noise_std=NOISE_STD, frames=n_frames,
                              grid=f"{nz}x{nx} z 5..45 mm x +-19.2 mm",
                              scale_convention="tau = s * (offsets + r)"))


def config2(seed: int = 2, n: int = 128, n_samples: int = 4096, n_frames: int = 1) -> Workload:
    """Plain 3-D volumetric das: 32x32 matrix array (pitch 0.3 mm), 16 diverging waves from
    virtual sources 10 mm behind the array on a 4x4 grid (3 mm lateral spacing), 4096 samples,
    n^3 grid over +-15 mm lateral and 5..50 mm depth, 50 scatterers.  Directivity mask via
    srcdir = (0,0,1,cos 30 deg); Hann angular window per (transmit, voxel) as weights."""
    rng = np.random.default_rng(seed)
    elem = matrix_array(32, 0.3e-3)
    g = (np.arange(4) - 1.5) * 3e-3
    vx, vy = np.meshgrid(g, g, indexing="ij")
    sources = np.stack([vx.ravel(), vy.ravel(), np.full(16, -10e-3)], axis=-1)
    half_ap = 0.5 * 32 * 0.3e-3
    # distance from each source to the aperture rectangle (reference utils/spatial.py:25-37)
    ddx = np.clip(np.abs(sources[:, 0]) - half_ap, 0, None)
    ddy = np.clip(np.abs(sources[:, 1]) - half_ap, 0, None)
    ref = np.sqrt(ddx ** 2 + ddy ** 2 + sources[:, 2] ** 2)
    P = 50
    pos = np.stack([rng.uniform(-15e-3, 15e-3, P), rng.uniform(-15e-3, 15e-3, P),
                    rng.uniform(5e-3, 50e-3, P)], -1)
    amp = rng.uniform(0.5, 1.0, P)
    x = synth_echoes(rng, n_frames, elem, lambda p: diverging_delay(p, sources, ref), 16,
                     n_samples, lambda b: (pos, amp))
    lat = np.linspace(-15e-3, 15e-3, n)
    dep = np.linspace(5e-3, 50e-3, n)
    zz, yy, xx = np.meshgrid(dep, lat, lat, indexing="ij")
    grid = np.stack([xx, yy, zz], axis=-1)                       # (n, n, n, 3)
    offsets = diverging_delay(grid, sources, ref).astype(np.float32)
    # Hann window of the angle between (voxel - source) and +z, f-number 1 (quickstart.md:66-81)
    vec = grid[None] - sources.reshape(16, 1, 1, 1, 3)
    cos_t = vec[..., 2] / np.linalg.norm(vec, axis=-1)
    ratio = np.tan(np.arccos(np.clip(cos_t, -1, 1)))
    weights = np.where(np.abs(ratio) <= 0.5, 0.5 + 0.5 * np.cos(2 * np.pi * ratio), 0.0)
    srcdir = np.tile(np.array([0.0, 0.0, 1.0, math.cos(math.radians(30.0))], np.float32),
                     (1024, 1))
    return Workload("config2_3d_das", x, (elem * KS).astype(np.float32),
                    (grid * KS).astype(np.float32), offsets, weights.astype(np.float32),
                    srcdir=srcdir,
                    meta=dict(seed=seed, n=n, samples=n_samples, frames=n_frames,
                              sources="4x4, 3 mm spacing, z=-10 mm",
                              weights="Hann(f# 1 angle window) per (transmit, voxel)",
                              srcdir="(0,0,1,cos30)"))


# ---------------------------------------------------------------------------------------
# config 5: batched sparse compounding over frame ensembles, mixed probes
# ---------------------------------------------------------------------------------------
def focused_delay(points_m: np.ndarray, foci_m: np.ndarray) -> np.ndarray:
    """One-way focused-transmit delay in samples, (T, *points): a virtual source at the focus
    f fed from the element below it, o = (f_x, f_y, 0): |f - o| + sign(z - f_z) |y - f|."""
    sh = (-1,) + (1,) * (points_m.ndim - 1) + (3,)
    f = foci_m.reshape(sh)
    d = np.linalg.norm(points_m[None] - f, axis=-1)
    sgn = np.where(points_m[None, ..., 2] >= f[..., 2], 1.0, -1.0)
    return (f[..., 2] + sgn * d) * KS


def nearest_foci(points_m: np.ndarray, foci_m: np.ndarray, k: int) -> np.ndarray:
    """int32 (k, *points): the k foci with the smallest lateral distance to each voxel."""
    lat = points_m[..., None, :2] - foci_m[:, :2]
    d = np.linalg.norm(lat, axis=-1)                                  # (*points, T)
    idx = np.argsort(d, axis=-1, kind="stable")[..., :k]
    return np.moveaxis(idx, -1, 0).astype(np.int32)


CONFIG5_PROBES = {
    # name: (elements, grid shape, lateral half width [m], depth range [m], foci layout)
    "linear64": dict(kind="linear", n=64, pitch=0.3e-3, grid=(128, 128), half=9.6e-3,
                     depth=(5e-3, 45e-3), focus_z=25e-3),
    "linear256": dict(kind="linear", n=256, pitch=0.2e-3, grid=(192, 192), half=25.6e-3,
                      depth=(5e-3, 45e-3), focus_z=25e-3),
    "matrix1024": dict(kind="matrix", n=32, pitch=0.3e-3, grid=(96, 96, 96), half=9.6e-3,
                       depth=(5e-3, 30e-3), focus_z=17.5e-3),
}


def config5_geometry(probe: str, n_tx: int = 64, k: int = 8, scale: float = 1.0):
    """Geometry of one config-5 probe: element positions, voxel grid, focused-transmit delays,
    the k-nearest-foci index table and weights 1/k.  `scale` shrinks the grid (tests).
    Returns (elem [m], grid [m], offsets [samples], weights, sparse_indices, foci [m])."""
    c = CONFIG5_PROBES[probe]
    if c["kind"] == "linear":
        elem = linear_array(c["n"], c["pitch"])
        nz, nx = (max(2, int(round(v * scale))) for v in c["grid"])
        grid = _grid2d(nz, nx, c["depth"], c["half"])
        fx = np.linspace(-c["half"], c["half"], n_tx)
        foci = np.stack([fx, np.zeros(n_tx), np.full(n_tx, c["focus_z"])], -1)
    else:
        elem = matrix_array(c["n"], c["pitch"])
        nd, ny, nx = (max(2, int(round(v * scale))) for v in c["grid"])
        lat_y = np.linspace(-c["half"], c["half"], ny)
        lat_x = np.linspace(-c["half"], c["half"], nx)
        dep = np.linspace(c["depth"][0], c["depth"][1], nd)
        zz, yy, xx = np.meshgrid(dep, lat_y, lat_x, indexing="ij")
        grid = np.stack([xx, yy, zz], axis=-1)
        m = int(round(math.sqrt(n_tx)))
        g = np.linspace(-c["half"], c["half"], m)
        fxx, fyy = np.meshgrid(g, g, indexing="ij")
        foci = np.stack([fxx.ravel(), fyy.ravel(), np.full(m * m, c["focus_z"])], -1)
    offsets = focused_delay(grid, foci)                               # (T, *grid)
    idx = nearest_foci(grid, foci, k)                                 # (k, *grid)
    offs_k = np.take_along_axis(offsets, idx.astype(np.int64), axis=0).astype(np.float32)
    weights = np.full(offs_k.shape, 1.0 / k, dtype=np.float32)
    return elem, grid, offs_k, weights, idx, foci


def config5(seed: int = 5, probe: str = "linear64", n_frames: int = 200, n_tx: int = 64,
            k: int = 8, n_samples: int = 2048, scale: float = 1.0, n_tissue: int = 40,
            n_moving: int = 20) -> Workload:
    """Batched sparse compounding for one probe: das_sparse over an n_frames ensemble, n_tx
    focused transmits, each voxel compounding its k nearest foci (int32 table (k, *grid),
    weights 1/k).  Scatterers: n_tissue static ones (amplitude x100) and n_moving in a 5 mm
    band at the focal depth moving 0.05 mm/frame laterally.  Echoes are synthesised on the
    host, so use this at reduced size (scale, frames, channels); the full-size benchmark
    draws channel data on the device (the timing does not depend on the sample values)."""
    rng = np.random.default_rng(seed)
    elem, grid, offs, weights, idx, foci = config5_geometry(probe, n_tx, k, scale)
    c = CONFIG5_PROBES[probe]
    half = c["half"] * 0.9
    d0, d1 = c["depth"]
    lat2 = c["kind"] == "matrix"
    tis = np.stack([rng.uniform(-half, half, n_tissue),
                    rng.uniform(-half, half, n_tissue) if lat2 else np.zeros(n_tissue),
                    rng.uniform(d0, d1, n_tissue)], -1)
    tis_amp = 100.0 * rng.uniform(0.5, 1.0, n_tissue)
    mv0 = np.stack([rng.uniform(-half, half, n_moving),
                    rng.uniform(-half, half, n_moving) if lat2 else np.zeros(n_moving),
                    rng.uniform(c["focus_z"] - 2.5e-3, c["focus_z"] + 2.5e-3, n_moving)], -1)
    mv_amp = rng.uniform(0.5, 1.0, n_moving)

    def scat(b):
        mv = mv0.copy()
        mv[:, 0] = (mv[:, 0] + b * 0.05e-3 + half) % (2 * half) - half
        return np.concatenate([tis, mv]), np.concatenate([tis_amp, mv_amp])

    x = synth_echoes(rng, n_frames, elem, lambda p: focused_delay(p, foci), n_tx, n_samples,
                     scat)
    return Workload(f"config5_{probe}", x, (elem * KS).astype(np.float32),
                    (grid * KS).astype(np.float32), offs, weights, sparse_indices=idx,
                    meta=dict(seed=seed, probe=probe, frames=n_frames, transmits=n_tx, k=k,
                              samples=n_samples, scale=scale,
                              transmit="focused, virtual source at the focus fed from the "
                                       "element below it",
                              selection="k nearest foci by lateral distance",
                              motion="0.05 mm/frame lateral, 5 mm band at the focal depth",
                              tissue="static, amplitude x100"))


def small_random(seed=0, B=2, C=6, T=3, S=96, N=(5, 7), K=None, complex_iq=False,
                 srcdir=False, sparse=False):
    """Small random problem with delays that land inside and outside [0, S-1)."""
    rng = np.random.default_rng(seed)
    K = K or T
    srcpos = np.stack([rng.uniform(-8, 8, C), rng.uniform(-2, 2, C), np.zeros(C)], -1)
    dst = np.stack([rng.uniform(-10, 10, N), rng.uniform(-3, 3, N), rng.uniform(4, 40, N)], -1)
    offsets = rng.uniform(-10, S * 0.6, (K,) + tuple(N))
    weights = rng.uniform(-1, 1, (K,) + tuple(N))
    shape = (B, C, T, S)
    x = rng.standard_normal(shape)
    if complex_iq:
        x = x + 1j * rng.standard_normal(shape)
    sd = None
    if srcdir:
        sd = np.tile(np.array([0, 0, 1, 0.6]), (C, 1))
        sd[::2, 3] = 0.2
    si = rng.integers(0, T, (K,) + tuple(N)).astype(np.int32) if sparse else None
    return Workload("small_random", x.astype(np.complex64 if complex_iq else np.float32),
                    srcpos.astype(np.float32), dst.astype(np.float32), offsets.astype(np.float32),
                    weights.astype(np.float32), srcdir=None if sd is None else sd.astype(np.float32),
                    wavenum=0.7 if complex_iq else 0.0, sparse_indices=si, meta=dict(seed=seed))
