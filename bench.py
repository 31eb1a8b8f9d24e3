"""Benchmark of the delay-and-sum hot path on one MI355X.

Workload (default): the differentiation configuration -- 2-D das forward + adjoint, 128-element
linear array, 11 plane waves (-15..+15 deg), 2048 samples, 16 frames, 256 x 256 voxel grid,
channel data held in FP16 (FP32 arithmetic).  One step = das forward, the L2-loss gradient
2 (out - target), and the adjoint (grad wrt channel data, transmit delays and the delay
scale).  Inputs are resident in HBM before timing starts.

Metric: Gvoxel*ch*tx/s, one unit U = one (voxel, channel, transmit, frame) interpolation,
counted once per forward+adjoint step.  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--storage fp16|bf16|fp32]
  python bench.py --profile-only N   # N steps + a 256 MiB copy, for rocprofv3 --pmc runs
  python bench.py --config 2         # volumetric 3-D das, the largest single call (3.4e10 U)
  python bench.py --config 5         # batched sparse compounding, three probes, one graph
  python bench.py --op truncate_rank # clutter filter (rank truncation) of a frame ensemble
  python bench.py --op interpolate   # scan conversion of a sector volume to a Cartesian grid
  python bench.py --op greens        # frequency-domain Green's function simulation

roofline.traffic (HBM bytes per launch of the dominant kernel) is read from
profiles/das_config4_<storage>_traffic.json, written by ffdas_amd.utils.pmc from two
rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of `bench.py --profile-only`.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxel·ch·tx/s DAS fwd+adjoint, fp16-in vs fp32 at rel-L2≤1e-3 vs FP64"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # the first tens of steps after start-up (or after the autotuner) run about 6 % slower
    # than the steady state on gfx950 (1.09-1.10 against 1.02-1.03 ms for the default
    # workload, three windows of 20 steps in a row): the timed steps start after 30
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--storage", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--cpu-sample-rows", type=int, default=256,
                    help="voxel rows (of 256) in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--config", type=int, default=4, choices=[2, 4, 5],
                    help="4: fwd+adjoint headline (default); 2: volumetric 3-D das (FP32, "
                         "1024 channels, 16 diverging waves, 128^3 voxels); 5: batched sparse "
                         "compounding")
    ap.add_argument("--op", default="das",
                    choices=["das", "truncate_rank", "interpolate", "greens"],
                    help="das (default): the delay-and-sum benchmarks; truncate_rank: the "
                         "clutter filter on a beamformed frame ensemble; interpolate: scan "
                         "conversion of beamformed frames; greens: frequency-domain "
                         "simulation")
    ap.add_argument("--c5-frames", type=int, default=200, help="config 5 ensemble size")
    ap.add_argument("--c5-probes", default="linear64,linear256,matrix1024")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the built-in launch heuristics instead of autotuned parameters")
    ap.add_argument("--tuning-cache", default=None, metavar="FILE",
                    help="load tuned launch parameters from FILE if it exists, else autotune "
                         "and save them there (a profiled run then times only the tuned "
                         "configuration)")
    ap.add_argument("--profile-only", type=int, default=0, metavar="N",
                    help="run N steps and a 256 MiB calibration copy, print no JSON")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="end an --op run with the 256 MiB calibration copy (for --pmc runs)")
    return ap.parse_args()


class Problem:
    def __init__(self, w, dev, storage):
        import torch
        self.w = w
        self.dev = dev
        x32 = torch.from_numpy(w.x).to(dev)
        self.x = {"fp16": x32.half(), "bf16": x32.bfloat16(), "fp32": x32}[storage]
        self.x32 = x32
        self.srcpos = torch.from_numpy(w.srcpos).to(dev)
        self.dstpos = torch.from_numpy(w.dstpos).to(dev)
        self.offsets = torch.from_numpy(w.offsets).to(dev)
        self.weights = torch.from_numpy(w.weights).to(dev)
        g = torch.Generator(device="cpu").manual_seed(1234)
        B = w.x.shape[0]
        self.target = torch.randn((B,) + tuple(w.dstpos.shape[:-1]), generator=g).to(dev)
        self.neg2_target = -2.0 * self.target      # for the loss gradient in one kernel
        self.scale = 1.0


def step(p, x, scale=1.0, algorithm=None):
    """One forward + adjoint (grad x, grad offsets, grad scale) of the L2 loss."""
    import torch
    import ffdas_amd
    alg = ffdas_amd.Algorithm.DEFAULT if algorithm is None else algorithm
    out = ffdas_amd.das(x, p.srcpos, p.dstpos, p.offsets, p.weights, scale=scale, algorithm=alg)
    # d/d out of sum (out - target)^2 = 2 out - 2 target, one fused kernel (scaling by 2 is
    # exact, so this equals 2 (out - target) bit for bit)
    g = torch.add(p.neg2_target, out, alpha=2.0)
    # x is unchanged since the forward: the adjoint may gather from the forward's converted
    # copy of x (the same contract autograd gives for its saved tensors)
    gx, go, _, gs = ffdas_amd.das_backward(x, p.srcpos, p.dstpos, p.offsets, p.weights, g,
                                           scale=scale, need_weights=False, algorithm=alg,
                                           reuse_x=True)
    return out, gx, go, gs


def tune(p, x):
    """Autotune the forward and adjoint launch parameters for this problem (kept in the
    library's tuning cache for the timed steps)."""
    import ffdas_amd
    return ffdas_amd.autotune(x, p.srcpos, p.dstpos, p.offsets, p.weights, reps=5)


def bwdx_frames_used(a, tuned):
    """Frames per grad_x chunk of the timed storage type: from the autotuner's result or the
    loaded tuning file; 16 (the library's default for real data here) when untuned."""
    fs = 0
    if tuned and "bwdx_frames" in tuned:
        fs = tuned["bwdx_frames"]
    elif tuned and "loaded" in tuned:
        tag = "|x=" + {"fp16": "f16", "bf16": "bf16", "fp32": "f32"}[a.storage] + "|"
        with open(tuned["loaded"]) as f:
            for line in f:
                if tag in line:
                    fs = int(line.split()[6])
    return fs if fs > 0 else 16


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def kernel_times(p, x, steps):
    """Average duration of each library kernel over `steps` forward+adjoint steps, from the
    library's own hipEvent pairs recorded around every launch on the stream it runs on.  The
    adjoint's grad_x scatter, which in the timed steps runs on a second stream beside the
    parameter adjoint (16-bit storage), is serialised here (FFDAS_BWD_CONCURRENT=0) so that
    each duration is the kernel's own rather than a share of the overlap."""
    import torch
    from ffdas_amd.utils import profiling
    old = os.environ.get("FFDAS_BWD_CONCURRENT")
    os.environ["FFDAS_BWD_CONCURRENT"] = "0"
    try:
        for _ in range(2):
            step(p, x)
        torch.cuda.synchronize()
        with profiling.profiled():
            for _ in range(steps):
                step(p, x)
            torch.cuda.synchronize()
    finally:
        if old is None:
            del os.environ["FFDAS_BWD_CONCURRENT"]
        else:
            os.environ["FFDAS_BWD_CONCURRENT"] = old
    return {k: v["avg_ms"] for k, v in profiling.kernel_stats().items()}


def accuracy(p):
    """rel-L2 against the FP64 oracle on frame 0: the forward output and the adjoint's
    grad_x for the loss gradient of that frame."""
    import oracle
    import ffdas_amd
    w = p.w
    x0 = p.x[:1].contiguous()
    out = ffdas_amd.das(x0, p.srcpos, p.dstpos, p.offsets, p.weights)
    ref = oracle.das(w.x[:1], w.srcpos, w.dstpos, w.offsets, w.weights, bits=64)
    g = 2.0 * (out - p.target[:1])
    gx = ffdas_amd.das_backward(x0, p.srcpos, p.dstpos, p.offsets, p.weights, g,
                                need_offsets=False, need_weights=False, need_scale=False)[0]
    gref = oracle.das_adjoint(w.x[:1], w.srcpos, w.dstpos, w.offsets, w.weights,
                              g.cpu().numpy(), bits=64)["grad_x"]
    return (oracle.rel_l2(out.cpu().numpy(), ref),
            oracle.rel_l2(gx.float().cpu().numpy(), gref))


def traffic_for(storage, kernel, summary=None):
    """HBM bytes per launch of `kernel` from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", summary or f"das_config4_{storage}_traffic.json")
    try:
        with open(path) as fh:
            k = json.load(fh)["kernels"].get(kernel + "_kernel")
        return None if k is None else round(k["traffic_bytes"])
    except (OSError, ValueError, KeyError):
        return None


def counters_for(kernel, summary):
    """Per-kernel counter figures (VALU busy, ...) from a committed counter summary
    (profiles/<summary>, written by `python -m ffdas_amd.utils.pmc counters`), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", summary)) as fh:
            return json.load(fh)["kernels"].get(kernel + "_kernel")
    except (OSError, ValueError, KeyError):
        return None


def valu_onchip(kernel, summary):
    """onchip object for a VALU-bound kernel: its VALU busy share from the counter summary
    (SQ_INSTS_VALU x 4 cycles over CUs x SIMDs x the kernel's cycles), against 1."""
    c = counters_for(kernel, summary)
    busy = None if c is None else c.get("valu_busy")
    return {"resource": "VALU issue", "unit": "busy share", "achieved": busy and round(busy, 3),
            "peak": 1.0, "frac": busy and round(busy, 3),
            "peak_source": f"profiles/{summary}: SQ_INSTS_VALU x 4 cycles / (256 CUs x 4 SIMDs "
                           f"x kernel cycles), one counter run of the same workload"}


def calibration_copy():
    """A 256 MiB device copy, twice: the FETCH/WRITE_SIZE counters' calibration."""
    import torch
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()


def profile_only(p, x, n):
    """A few steps and a 256 MiB device copy (the FETCH/WRITE_SIZE calibration)."""
    for _ in range(n):
        step(p, x)
    calibration_copy()
    log(f"profile-only: {n} steps and the calibration copy done")


def lds_peak():
    """Conflict-free 64-bit LDS atomic rate of this device [lane-ops/clk/CU], measured live
    by the calibration probe (ffdas_amd/calib/lds_rate.hip: every lane of 8 waves per
    workgroup adds at its own bank pair), or None when the probe library is missing."""
    try:
        from ffdas_amd.utils import calib
        return calib.lds_atomic_rate("u64")
    except (ImportError, RuntimeError, OSError) as e:
        log(f"LDS calibration probe unavailable: {e}")
        return None


def roofline(a, p, kt, U, tuned):
    """Roofline object of the dominant kernel.  frac = HBM bytes per launch from the counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated) / live kernel time / 8 TB/s.  The
    algorithmic bytes (gathered samples or FP32 scatter-adds per unit) are reported apart:
    L1/L2 and LDS serve them, so they do not bound against HBM.  onchip: the resource that
    binds the kernel, against its measured peak."""
    esz = p.x.element_size()
    bytes_per_u = {"das_fwd_fi": 2 * esz, "das_fwd": 2 * esz, "das_bwd_param_fi": 2 * esz,
                   "das_bwd_param": 2 * esz, "das_bwd_x": 2 * 4, "das_bwd_x_plain": 2 * 4}
    dom = max((k for k in kt if k in bytes_per_u), key=kt.get)
    t = kt[dom] * 1e-3
    traffic = traffic_for(a.storage, dom)
    hbm = None if traffic is None else traffic / t / 1e9
    alg = U * bytes_per_u[dom]
    cus, ghz = 256, 2.4
    if dom == "das_bwd_x":
        # per unit: one 64-bit LDS atomic (a frame pair at each of the two samples) plus one
        # 32-bit count per tap and frame chunk of FS frames
        fs = bwdx_frames_used(a, tuned)
        ops = U * (1.0 + 1.0 / fs)
        cal = lds_peak()
        if cal is not None:
            cus, ghz = cal["cus"], cal["clock_ghz"]
        peak = round(cal["lane_ops_per_clk_cu"], 2) if cal else 10.05
        onchip = {"resource": "LDS 64-bit atomics", "unit": "lane-ops/clk/CU",
                  "achieved": round(ops / t / (ghz * 1e9 * cus), 2), "peak": peak,
                  "peak_source": "measured now: ffdas_calib_lds_atomics, distinct bank pairs"
                  if cal else "ffdas_calib_lds_atomics on gfx950 (earlier run)"}
    else:
        onchip = {"resource": "vector L1", "unit": "B/clk/CU",
                  "achieved": round(alg / t / (ghz * 1e9 * cus), 1), "peak": 64.0,
                  "peak_source": "gfx950 vector L1 line rate, 64 B/clk/CU"}
    onchip["frac"] = round(onchip["achieved"] / onchip["peak"], 3)
    return {"bound": "hbm", "kernel": dom,
            "achieved": None if hbm is None else round(hbm, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": None if hbm is None else round(hbm / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": f"profiles/das_config4_{a.storage}_traffic.json",
            "algorithmic": {"bytes_per_unit": bytes_per_u[dom], "units_per_launch": U,
                            "bytes_per_launch": alg, "GBps": round(alg / t / 1e9, 1),
                            "vs_hbm_peak": round(alg / t / 1e9 / HBM_PEAK_GBS, 3)},
            "onchip": onchip,
            "kernel_ms": {k: round(v, 4) for k, v in kt.items()}}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def cpu_baseline(w, rows, single_rows=16):
    """The FP32 build of the oracle on the host cores: forward + adjoint (grad x, offsets,
    scale) over a bounded sample (the first `rows` voxel rows of the grid, all frames,
    channels and transmits) on all host threads, median of 5; and the same on one thread over
    the first `single_rows` rows, median of 3."""
    import oracle
    nthreads = oracle.hardware_threads()

    def run(nrows, threads, reps):
        dst = np.ascontiguousarray(w.dstpos[:nrows])
        ofs = np.ascontiguousarray(w.offsets[:, :nrows])
        wt = np.ascontiguousarray(w.weights[:, :nrows])
        B, C, T, S = w.x.shape
        U = B * C * T * dst.shape[0] * dst.shape[1]
        g = np.random.default_rng(0).standard_normal((B,) + dst.shape[:-1]).astype(np.float32)
        oracle.das(w.x, w.srcpos, dst, ofs, wt, bits=32, threads=threads)     # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            oracle.das(w.x, w.srcpos, dst, ofs, wt, bits=32, threads=threads)
            oracle.das_adjoint(w.x, w.srcpos, dst, ofs, wt, g, bits=32, threads=threads)
            times.append(time.perf_counter() - t0)
        return U, U / float(np.median(times)) / 1e9

    U, v = run(rows, nthreads, 5)
    U1, v1 = run(single_rows, 1, 3)
    return {"value": v, "unit": "Gvoxel·ch·tx/s", "cores": nthreads, "kind": "port",
            "cpu_model": cpu_model(),
            "single_core": {"value": v1, "cores": 1,
                            "sample": f"first {single_rows} of 256 voxel rows ({U1:.3g} U), "
                                      f"median of 3"},
            "sample": f"FP32 oracle fwd+adjoint on {rows}x256 of the 256x256 voxels, all 16 "
                      f"frames/128 ch/11 tx ({U:.3g} U), median of 5, {nthreads} threads"}


def bench_config2(a):
    """Plain volumetric das (FP32): 32x32 matrix array (1024 channels), 16 diverging waves
    from virtual sources 10 mm behind the array, 4096 samples, 128^3 voxels over +-15 mm x
    5-50 mm, directivity cone 30 deg and Hann angular weights: the largest single call of the
    benchmark set (3.436e10 units).  One step = one forward call; synthesised echoes of 50
    scatterers."""
    import torch
    import ffdas_amd
    from ffdas_amd import workloads as W
    dev = torch.device("cuda:0")
    t0 = time.time()
    w = W.config2()
    log(f"config 2 generated in {time.time() - t0:.1f}s: x {w.x.shape}, grid "
        f"{w.dstpos.shape[:-1]}, U {w.units:.4g}")
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    x, srcpos, dstpos = t(w.x), t(w.srcpos), t(w.dstpos)
    offsets, weights, srcdir = t(w.offsets), t(w.weights), t(w.srcdir)
    run = lambda: ffdas_amd.das(x, srcpos, dstpos, offsets, weights, srcdir=srcdir)
    if a.profile_only:
        for _ in range(a.profile_only):
            run()
        calibration_copy()
        log(f"profile-only: {a.profile_only} calls and the calibration copy done")
        return
    steps = max(2, min(a.steps, 10))
    ms, wall = time_steps(run, steps, max(1, a.warmup))
    U = w.units
    value = U / (ms * 1e-3) / 1e9
    from ffdas_amd.utils import profiling
    with profiling.profiled():
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
    kt = {k: v["avg_ms"] for k, v in profiling.kernel_stats().items()}
    log(f"config 2: {ms:.2f} ms per call, {value:.1f} G U/s; kernels {kt}")
    dom = max(kt, key=kt.get)
    tk = kt[dom] * 1e-3
    traffic = traffic_for("fp32", dom, summary="das_config2_fp32_traffic.json")
    hbm = None if traffic is None else traffic / tk / 1e9
    alg = U * 8
    # the transmit tables (offsets, weights: 16 x 128^3 FP32 each) are read once per call
    tables = 2 * w.offsets.size * 4
    roof = {"bound": "hbm", "kernel": dom,
            "achieved": None if hbm is None else round(hbm, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": None if hbm is None else round(hbm / HBM_PEAK_GBS, 4),
            "traffic": traffic, "traffic_source": "profiles/das_config2_fp32_traffic.json",
            "algorithmic": {"bytes_per_unit": 8, "units_per_launch": U, "bytes_per_launch": alg,
                            "GBps": round(alg / tk / 1e9, 1),
                            "vs_hbm_peak": round(alg / tk / 1e9 / HBM_PEAK_GBS, 3)},
            "transmit_table_bytes": tables,
            "onchip": valu_onchip(dom, "das_config2_fp32_counters.json"),
            "kernel_ms": {k: round(v, 4) for k, v in kt.items()}}
    cpu = None
    if not a.no_cpu:
        import oracle
        nz = 4            # a slab of 4 of the 128 depth planes, all lateral voxels
        dst = np.ascontiguousarray(w.dstpos[:nz])
        ofs = np.ascontiguousarray(w.offsets[:, :nz])
        wts = np.ascontiguousarray(w.weights[:, :nz])
        n = int(np.prod(dst.shape[:-1])) * w.x.shape[1] * w.offsets.shape[0]
        oracle.das(w.x, w.srcpos, dst[:1, :8], ofs[:, :1, :8], wts[:, :1, :8],
                   srcdir=w.srcdir, bits=32)
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            oracle.das(w.x, w.srcpos, dst, ofs, wts, srcdir=w.srcdir, bits=32)
            ts.append(time.perf_counter() - t1)
        d1, o1, w1 = dst[:1, :16], ofs[:, :1, :16], wts[:, :1, :16]
        n1 = int(np.prod(d1.shape[:-1])) * w.x.shape[1] * w.offsets.shape[0]
        t1 = time.perf_counter()
        oracle.das(w.x, w.srcpos, d1, o1, w1, srcdir=w.srcdir, bits=32, threads=1)
        v1 = n1 / (time.perf_counter() - t1) / 1e9
        cpu = {"value": n / float(np.median(ts)) / 1e9, "unit": "Gvoxel·ch·tx/s",
               "cores": oracle.hardware_threads(), "kind": "port", "cpu_model": cpu_model(),
               "single_core": {"value": v1, "cores": 1,
                               "sample": f"16 voxels x 1024 ch x 16 tx ({n1:.3g} U), one run"},
               "sample": f"FP32 oracle forward on {nz} of the 128 depth planes (all 128x128 "
                         f"lateral voxels, 1024 ch, 16 tx: {n:.3g} U), median of 3"}
        log(f"cpu baseline: {cpu}")
    line = {"metric": "Gvoxel·ch·tx/s volumetric das forward (FP32)", "value": round(value, 2),
            "unit": "Gvoxel·ch·tx/s", "n_gpus": 1, "steps": steps, "warmup": max(1, a.warmup),
            "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "3-D das: 32x32 matrix array (1024 ch), 16 diverging waves, "
                                   "4096 samples, 128^3 voxels, directivity 30 deg, Hann "
                                   "weights", "storage": "fp32", "units_per_step": U,
                       "wall_s": round(wall, 3)},
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def bench_config5(a):
    """das_sparse over frame ensembles for mixed probes in one batched launch, captured in a
    HIP graph and replayed (FP16 storage, FP32 accumulation).  Channel data are drawn on the
    device (standard normal, FP16): the gather timing does not depend on sample values, and
    the echo synthesis of the full ensembles (54 GB for the matrix probe) would dominate the
    run; parity of this path is covered at reduced size with synthesised echoes."""
    import torch
    import ffdas_amd
    from ffdas_amd import workloads as W
    dev = torch.device("cuda:0")
    names = [n for n in a.c5_probes.split(",") if n]
    probs, U, geo_s = [], 0, time.time()
    for name in names:
        elem, grid, offs, wts, idx, _ = W.config5_geometry(name)
        C = elem.shape[0]
        B, T, S = a.c5_frames, 64, 2048
        g = torch.Generator(device=dev).manual_seed(5)
        x = torch.randn((B, C, T, S), dtype=torch.float16, device=dev, generator=g)
        t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
        probs.append(dict(x=x, srcpos=t((elem * W.KS).astype(np.float32)),
                          dstpos=t((grid * W.KS).astype(np.float32)), offsets=t(offs),
                          weights=t(wts), sparse_indices=t(idx)))
        U += int(np.prod(grid.shape[:-1])) * C * idx.shape[0] * B
        log(f"{name}: x {tuple(x.shape)} fp16 {x.numel() * 2 / 1e9:.1f} GB, grid "
            f"{grid.shape[:-1]}, k={idx.shape[0]}")
    log(f"geometry and data in {time.time() - geo_s:.1f}s, U/ensemble {U:.4g}")
    run = lambda: ffdas_amd.das_batched(probs)
    if a.profile_only:
        for _ in range(a.profile_only):
            run()
        calibration_copy()
        log(f"profile-only: {a.profile_only} calls and the calibration copy done")
        return
    ms_eager, _ = time_steps(run, a.steps, a.warmup)
    cap = ffdas_amd.CapturedCall(run)
    ms_graph, wall = time_steps(cap.replay, a.steps, a.warmup)
    from ffdas_amd.utils import profiling
    with profiling.profiled():
        for _ in range(max(1, a.steps)):
            run()
        torch.cuda.synchronize()
    kt = {k: v["avg_ms"] for k, v in profiling.kernel_stats().items()}
    value = U / (ms_graph * 1e-3) / 1e9
    log(f"eager {ms_eager:.2f} ms, graph {ms_graph:.2f} ms per ensemble: {value:.1f} G U/s; "
        f"kernels {kt}")
    dom = "das_fwd_fi_batched"
    tk = kt[dom] * 1e-3
    traffic = traffic_for("fp16", dom, summary="das_config5_fp16_traffic.json")
    hbm = None if traffic is None else traffic / tk / 1e9
    alg = U * 4
    cpu = None
    if not a.no_cpu:
        import oracle
        w = W.config5(probe="linear64", n_frames=16, scale=1.0)
        n = w.units
        oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights,
                   sparse_indices=w.sparse_indices, bits=32)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.das(w.x, w.srcpos, w.dstpos, w.offsets, w.weights,
                       sparse_indices=w.sparse_indices, bits=32)
            ts.append(time.perf_counter() - t0)
        cpu = {"value": n / float(np.median(ts)) / 1e9, "unit": "Gvoxel·ch·tx/s",
               "cores": oracle.hardware_threads(), "kind": "port",
               "sample": f"FP32 oracle das_sparse, 64-element probe, 128x128 voxels, 64 "
                         f"transmits k=8, 16 frames ({n:.3g} U), median of 3"}
    line = {"metric": "Gvoxel·ch·tx/s batched das_sparse (mixed probes, one graph)",
            "value": round(value, 2), "unit": "Gvoxel·ch·tx/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_graph, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (device-drawn FP16 noise)",
            "config": {"workload": "batched sparse compounding: " + ",".join(names) +
                                   f", {a.c5_frames} frames, 64 focused transmits, k=8, "
                                   "2048 samples", "storage": "fp16", "units_per_step": U,
                       "parity": "timed on device-drawn noise; the same batched path is "
                                 "checked against the FP64 oracle with synthesised echoes at "
                                 "<= 5 frames (tests/test_config5_gpu.py) and on 2 frames of "
                                 "the full probe geometries (tests/test_fullsize_gpu.py)",
                       "eager_ms": round(ms_eager, 3), "wall_s": round(wall, 3)},
            "roofline": {"bound": "hbm", "kernel": dom,
                         "achieved": None if hbm is None else round(hbm, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": None if hbm is None else round(hbm / HBM_PEAK_GBS, 4),
                         "traffic": traffic,
                         "traffic_source": "profiles/das_config5_fp16_traffic.json",
                         "algorithmic": {"bytes_per_unit": 4, "units_per_launch": U,
                                         "bytes_per_launch": alg,
                                         "GBps": round(alg / tk / 1e9, 1),
                                         "vs_hbm_peak": round(alg / tk / 1e9 / HBM_PEAK_GBS, 3)},
                         "onchip": valu_onchip(dom, "das_config5_fp16_counters.json"),
                         "kernel_ms": {k: round(v, 3) for k, v in kt.items()}},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def sector_grid(nr, nt, nphi, r0=0.05, r1=1.0, half=0.6):
    """Vertices (nr, nt, nphi, 3) of a volumetric sector scan: radius r0..r1 along z, two
    steering angles in [-half, half] rad (FP32)."""
    import numpy as np
    r = np.linspace(r0, r1, nr)
    th = np.linspace(-half, half, nt)
    ph = np.linspace(-half, half, nphi)
    R, TH, PH = np.meshgrid(r, th, ph, indexing="ij")
    return np.stack([R * np.sin(TH) * np.cos(PH), R * np.sin(PH), R * np.cos(TH) * np.cos(PH)],
                    -1).astype(np.float32)


def bench_interpolate(a):
    """Scan conversion after beamforming: interpolate 16 frames of a 512 x 96 x 96 sector
    volume (radius x two steering angles, FP32 envelope) onto a 256^3 Cartesian grid over its
    bounding box, with the plan built and the query points located once (preprocess), as a
    per-frame display loop runs it.  A step is one evaluation of all 16 frames.  Data:
    standard normal on the device."""
    import ctypes
    import numpy as np
    import torch
    import ffdas_amd
    from ffdas_amd._lib import lib
    dev = torch.device("cuda:0")
    L = lib()

    def kernel_ms(name):
        t, c = ctypes.c_double(), ctypes.c_int64()
        L.ffdas_profile_query(name.encode(), ctypes.byref(t), ctypes.byref(c))
        return t.value / max(c.value, 1) if c.value else 0.0

    B, nr, nt, nphi, nq1 = 16, 512, 96, 96, 256
    g = sector_grid(nr, nt, nphi)
    flat = g.reshape(-1, 3)
    ax = [np.linspace(flat[:, k].min(), flat[:, k].max(), nq1, dtype=np.float32) for k in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    q = torch.from_numpy(np.stack([X, Y, Z], -1).astype(np.float32)).to(dev)
    gd = torch.from_numpy(g).to(dev)
    gen = torch.Generator(device=dev).manual_seed(6)
    x = torch.randn((B, nr, nt, nphi), device=dev, generator=gen)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    interp = ffdas_amd.Interpolator(gd, mode="linear")
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    L.ffdas_profile_reset()
    L.ffdas_profile_enable(1)
    interp.preprocess(q)
    torch.cuda.synchronize()
    L.ffdas_profile_enable(0)
    locate_ms = kernel_ms("interp_locate")
    out = torch.empty((B,) + tuple(q.shape[:-1]), device=dev)
    steps = max(5, a.steps)
    ms, wall = time_steps(lambda: interp(x, q, out=out), steps, a.warmup)
    L.ffdas_profile_reset()
    L.ffdas_profile_enable(1)
    for _ in range(max(3, steps // 4)):
        interp(x, q, out=out)
    torch.cuda.synchronize()
    L.ffdas_profile_enable(0)
    eval_ms = kernel_ms("interp_eval")
    nq, npts = nq1 ** 3, nr * nt * nphi
    located = float(torch.isfinite(interp(x[:1], q, fill=float("nan"))).float().mean())
    value = B * nq / (ms * 1e-3) / 1e9
    log(f"interpolate {B} x {nr}x{nt}x{nphi} -> {nq1}^3: {ms:.3f} ms/step ({value:.1f} G/s), "
        f"eval kernel {eval_ms:.3f} ms, plan {plan_ms:.1f} ms, locate {locate_ms:.3f} ms, "
        f"located {located:.3f}")
    # algorithmic bytes of one evaluation: every frame of x once, every output once, the
    # lookup (4 ids + 4 weights per point) once per chunk of 16 frames (FP32 data)
    nbytes = B * (nq + npts) * 4 + -(-B // 16) * nq * 32
    roof = {"kernel": "interp_eval", "bound": "hbm", "ms": round(eval_ms, 4),
            "achieved": round(nbytes / (eval_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "bytes_per_launch": nbytes,
            "traffic": traffic_for(None, "interp_linear", "interpolate_scan_conversion_traffic.json")}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    cpu = None
    if not a.no_cpu:
        import oracle.interpolation as OI
        # the oracle end to end (plan, location, evaluation) on a 64 x 24 x 24 sector and
        # 4000 of its box's points, 16 frames
        gs = sector_grid(64, 24, 24)
        fs = gs.reshape(-1, 3)
        rng = np.random.default_rng(7)
        qs = rng.uniform(fs.min(0), fs.max(0), (4000, 3)).astype(np.float32)
        xs = rng.standard_normal((B,) + gs.shape[:3]).astype(np.float32)
        t0 = time.perf_counter()
        OI.interpolate(gs, xs, qs, mode="linear")
        secs = time.perf_counter() - t0
        cpu = {"value": float(f"{B * qs.shape[0] / secs / 1e9:.3g}"), "unit": "Gsample/s",
               "cores": 1, "kind": "port",
               "sample": "oracle interpolate (numpy, one thread) of 16 frames of a 64 x 24 x 24 "
                         "sector at 4000 points, plan and location included"}
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "Gsample/s interpolate (query points x frames per second, preprocessed plan)",
        "value": round(value, 3), "unit": "Gsample/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"scan conversion: {B} frames of a {nr} x {nt} x {nphi} sector "
                               f"volume (FP32) to {nq1}^3 Cartesian points, linear",
                   "plan_ms": round(plan_ms, 2), "locate_ms": round(locate_ms, 3),
                   "located_fraction": round(located, 4), "wall_s": round(wall, 3)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_greens(a):
    """Frequency-domain simulation step: greens from a 128-element linear array (0.3 mm
    pitch) to 32768 points of a 128 x 256 image plane, 512 frequency bins around 5 MHz, 8
    transmit events (batch), complex64.  A step is one call.  Data: standard normal field
    on the device."""
    import ctypes
    import math
    import numpy as np
    import torch
    import ffdas_amd
    from ffdas_amd._lib import lib
    dev = torch.device("cuda:0")
    L = lib()
    B, ns, nf = 8, 128, 512
    nz, nx = 256, 128
    src = np.zeros((ns, 3), np.float32)
    src[:, 0] = (np.arange(ns) - (ns - 1) / 2) * 0.3e-3
    zz, xx = np.meshgrid(np.linspace(5e-3, 45e-3, nz), np.linspace(-19e-3, 19e-3, nx),
                         indexing="ij")
    dst = np.stack([xx, np.zeros_like(xx), zz], -1).astype(np.float32)
    freqs = 5e6 + (np.arange(nf) - nf / 2) * (8e6 / nf)
    k = (-2 * math.pi * freqs / 1540.0).astype(np.float32)
    gen = torch.Generator(device=dev).manual_seed(8)
    x = torch.complex(torch.randn((B, ns, nf), device=dev, generator=gen),
                      torch.randn((B, ns, nf), device=dev, generator=gen))
    srcd, dstd, kd = (torch.from_numpy(v).to(dev) for v in (src, dst, k))
    out = torch.empty((B, nz, nx, nf), dtype=torch.complex64, device=dev)
    steps = max(5, a.steps)
    ms, wall = time_steps(lambda: ffdas_amd.greens(srcd, kd, x, dstd, out=out), steps, a.warmup)
    t, c = ctypes.c_double(), ctypes.c_int64()
    L.ffdas_profile_reset()
    L.ffdas_profile_enable(1)
    for _ in range(3):
        ffdas_amd.greens(srcd, kd, x, dstd, out=out)
    torch.cuda.synchronize()
    L.ffdas_profile_enable(0)
    L.ffdas_profile_query(b"greens", ctypes.byref(t), ctypes.byref(c))
    kms = t.value / max(c.value, 1)
    terms = B * nz * nx * ns * nf
    value = terms / (ms * 1e-3) / 1e9
    log(f"greens {B} x {ns} sources -> {nz}x{nx} points x {nf} freqs: {ms:.3f} ms/step, "
        f"{value:.1f} G terms/s, kernel {kms:.3f} ms")
    # 8 or more events run the complex multiply-adds on the FP32 MFMA: 8 algorithmic flops a
    # term against its dense peak (the 16-event tile is half used at 8 events, and every
    # (target, source, frequency) also costs a sincos on the VALU, overlapped with the MFMA)
    roof = {"kernel": "greens", "bound": "mfma", "ms": round(kms, 4),
            "achieved": round(terms * 8 / (kms * 1e-3) / 1e12, 2), "peak": 157.3,
            "unit": "TFLOP/s",
            "traffic": traffic_for(None, "greens_mfma", "greens_array_simulation_traffic.json")}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    cpu = None
    if not a.no_cpu:
        import oracle.greens as OG
        # the FP64 oracle (numpy) on 512 of the points, all sources, frequencies and events
        xs = x.cpu().numpy()
        t0 = time.perf_counter()
        OG.greens(src, k, xs, dst.reshape(-1, 3)[:512])
        secs = time.perf_counter() - t0
        cpu = {"value": round(B * 512 * ns * nf / secs / 1e9, 4), "unit": "Gterm/s",
               "cores": 1, "kind": "port",
               "sample": "oracle greens (numpy FP64) at 512 of the 32768 points"}
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "Gterm/s greens (source x target x frequency x event terms per second)",
        "value": round(value, 2), "unit": "Gterm/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"greens: {ns}-element array to {nz} x {nx} points, {nf} "
                               f"frequencies, {B} events, complex64",
                   "wall_s": round(wall, 3)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_truncate_rank(a):
    """The clutter filter after beamforming: truncate_rank of a frame ensemble, removing the
    strongest singular components (tissue).  Workload: the reference's clutter-filter example
    size, 64 frames of a 64^3 complex64 volume, start = 5; also timed, and reported in
    `config.also`: 200 frames of a 96^3 FP32 volume (the volumetric sparse-compounding
    ensemble), start = 3.  Data: standard normal on the device.  Each call synchronises once
    (after the eigendecomposition), so calls are timed by wall clock after a device sync;
    kernel durations come from the library's hipEvent profiling on the launch stream."""
    import ctypes
    import torch
    import ffdas_amd
    from ffdas_amd._lib import lib
    dev = torch.device("cuda:0")
    L = lib()

    def kernel_ms(name):
        t, c = ctypes.c_double(), ctypes.c_int64()
        L.ffdas_profile_query(name.encode(), ctypes.byref(t), ctypes.byref(c))
        return t.value / max(c.value, 1) if c.value else 0.0

    def run(shape, dtype, start, steps, warmup):
        g = torch.Generator(device=dev).manual_seed(5)
        x = torch.randn(shape, dtype=dtype, device=dev, generator=g)
        out = torch.empty_like(x)
        for _ in range(warmup):
            ffdas_amd.truncate_rank(x, start, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ffdas_amd.truncate_rank(x, start, out=out)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        L.ffdas_profile_reset()
        L.ffdas_profile_enable(1)
        for _ in range(max(3, steps // 4)):
            ffdas_amd.truncate_rank(x, start, out=out)
        torch.cuda.synchronize()
        L.ffdas_profile_enable(0)
        names = ("rank_gram", "rank_gram_reduce", "rank_jacobi", "rank_tridiag",
                 "rank_tridiag_eigvec", "rank_backtransform", "rank_project")
        kt = {k: kernel_ms(k) for k in names}
        return x, ms, {k: v for k, v in kt.items() if v > 0.0}

    steps = max(5, a.steps)
    m, n = 64, 64 ** 3
    x, ms, kt = run((m, 64, 64, 64), torch.complex64, 5, steps, a.warmup)
    value = m * n / (ms * 1e-3) / 1e9
    log(f"truncate_rank (64, 64^3) complex64, start 5: {ms:.3f} ms/call, kernels {kt}")
    xb = x.numel() * x.element_size()
    p = m
    # Gram: the p(p+1)/2 distinct entries, 8 real flops per complex multiply-add
    gram_flops = p * (p + 1) // 2 * n * 8
    gram = {"kernel": "rank_gram", "bound": "mfma", "ms": round(kt["rank_gram"], 4),
            "achieved": round(gram_flops / (kt["rank_gram"] * 1e-3) / 1e12, 2),
            "peak": 157.3, "unit": "TFLOP/s"}
    gram["frac"] = round(gram["achieved"] / gram["peak"], 4)
    proj = {"kernel": "rank_project", "bound": "hbm", "ms": round(kt["rank_project"], 4),
            "achieved": round(2 * xb / (kt["rank_project"] * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_launch": 2 * xb}
    proj["frac"] = round(proj["achieved"] / proj["peak"], 4)

    m2, n2 = 200, 96 ** 3
    x2, ms2, kt2 = run((m2, 96, 96, 96), torch.float32, 3, max(3, steps // 2), 1)
    log(f"truncate_rank (200, 96^3) float32, start 3: {ms2:.3f} ms/call, kernels {kt2}")
    del x2

    cpu = None
    if not a.no_cpu:
        import oracle.rank as R
        # one numpy (LAPACK) SVD filter of a quarter of the volume (64 x 65536 complex64)
        xs = x.reshape(m, n)[:, : n // 4].cpu().numpy()
        _, secs = R.svd_filter_timed(xs, 5)
        cpu = {"value": round(m * (n // 4) / secs / 1e9, 4), "unit": "Gsample/s",
               "cores": os.cpu_count(), "kind": "port",
               "sample": "numpy SVD filter (LAPACK, all host threads) of 64 frames x 65536 "
                         "voxels complex64, start 5"}
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "Gsample/s truncate_rank (frame x voxel elements filtered per second)",
        "value": round(value, 3), "unit": "Gsample/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": "truncate_rank clutter filter: 64 frames x 64^3 voxels "
                               "complex64, start 5",
                   "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
                   "also": {"workload": "200 frames x 96^3 voxels float32, start 3",
                            "ms_per_call": round(ms2, 4),
                            "value": round(m2 * n2 / (ms2 * 1e-3) / 1e9, 3),
                            "kernel_ms": {k: round(v, 4) for k, v in kt2.items()}}},
        "roofline": proj, "roofline_gram": gram, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only: sharding a call "
                                   "across GPUs is not built", "n_gpus": a.gpus}))
        sys.exit(2)
    ops = {"truncate_rank": bench_truncate_rank, "interpolate": bench_interpolate,
           "greens": bench_greens}
    if a.op in ops:
        ops[a.op](a)
        if a.calibration_copy:
            calibration_copy()
        return
    if a.config == 5:
        bench_config5(a)
        return
    if a.config == 2:
        bench_config2(a)
        return
    import torch
    import ffdas_amd
    from ffdas_amd import workloads as W

    dev = torch.device("cuda:0")
    t0 = time.time()
    w = W.config4()
    log(f"workload generated in {time.time() - t0:.1f}s: x {w.x.shape}, grid "
        f"{w.dstpos.shape[:-1]}, U/step {w.units:.4g}")
    p = Problem(w, dev, a.storage)
    U = w.units
    tuned = None
    if a.tuning_cache and os.path.exists(a.tuning_cache):
        ffdas_amd.load_tuning(a.tuning_cache)
        tuned = {"loaded": a.tuning_cache}
        log(f"tuning loaded from {a.tuning_cache}")
    elif not a.no_autotune:
        t0 = time.time()
        tuned = tune(p, p.x)
        if a.storage != "fp32":
            tune(p, p.x32)
        log(f"autotuned in {time.time() - t0:.1f}s: {tuned}")
        if a.tuning_cache:
            ffdas_amd.save_tuning(a.tuning_cache)
    if a.profile_only:
        profile_only(p, p.x, a.profile_only)
        return

    ms, wall = time_steps(lambda: step(p, p.x), a.steps, a.warmup)
    value = U / (ms * 1e-3) / 1e9
    log(f"{a.storage} fwd+adj: {ms:.3f} ms/step, {value:.1f} G U/s")

    # plain FP32 kernels (natural layout, ALG1) for the headline ratio, and the tuned FP32 path
    ms32, _ = time_steps(lambda: step(p, p.x32, algorithm=ffdas_amd.Algorithm.ALG1),
                         max(3, a.steps // 2), a.warmup)
    v32 = U / (ms32 * 1e-3) / 1e9
    log(f"fp32 plain (ALG1) fwd+adj: {ms32:.3f} ms/step, {v32:.1f} G U/s, ratio {value / v32:.2f}x")
    ms32o, _ = time_steps(lambda: step(p, p.x32), max(3, a.steps // 2), a.warmup)
    v32o = U / (ms32o * 1e-3) / 1e9
    log(f"fp32 tuned (DEFAULT) fwd+adj: {ms32o:.3f} ms/step, {v32o:.1f} G U/s")

    kt = kernel_times(p, p.x, max(3, a.steps // 2))
    log("kernel ms: " + ", ".join(f"{k}={v:.3f}" for k, v in kt.items()))
    kt32 = kernel_times(p, p.x32, max(3, a.steps // 2))
    log("fp32 tuned kernel ms: " + ", ".join(f"{k}={v:.3f}" for k, v in kt32.items()))
    roof = roofline(a, p, kt, U, tuned)

    # the CPU leg: the oracle timed on the host cores (FP32 build), and its FP64 build as the
    # checker of frame 0
    acc = (None, None)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(w, a.cpu_sample_rows)
        log(f"cpu baseline: {cpu}")
        if not a.no_accuracy:
            acc = accuracy(p)
            log(f"rel-L2 vs FP64 oracle (frame 0): out {acc[0]:.3e}, grad_x {acc[1]:.3e}")

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Gvoxel·ch·tx/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "2-D das fwd+adjoint: 128-ch linear array, 11 plane waves, "
                               "2048 samples, 16 frames, 256x256 voxels",
                   "storage": a.storage, "units_per_step": U,
                   "fp32_plain_value": round(v32, 2),
                   "fp32_tuned_value": round(v32o, 2),
                   "ratio_vs_fp32_plain": round(value / v32, 3),
                   "ratio_vs_fp32_tuned": round(value / v32o, 3),
                   "rel_l2_vs_fp64": acc[0], "rel_l2_grad_x_vs_fp64": acc[1],
                   "kernel_ms_note": "per-kernel durations from a pass with the adjoint's "
                                     "grad_x scatter and parameter adjoint serialised; the "
                                     "timed steps run them side by side (16-bit storage)",
                   "autotuned": None if tuned is None else
                   {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tuned.items()},
                   "wall_s": round(wall, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line["config"]["fp32_tuned_kernel_ms"] = {k: round(v, 4) for k, v in kt32.items()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
