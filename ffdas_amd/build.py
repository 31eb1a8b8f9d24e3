"""Build libffdas.so (the C-ABI library) in-tree with hipcc for gfx950.

    python -m ffdas_amd.build          # or ffdas_amd.build.build()

Host sources (*.cpp) and device sources (*.hip) are compiled in parallel into objects under
ffdas_amd/_build/ and linked into ffdas_amd/libffdas.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libffdas.so"
# measurement probe (LDS atomic rate), a library of its own: not part of the C-ABI
CALIB_SRC = PKG / "calib" / "lds_rate.hip"
CALIB_LIB = PKG / "calib" / "libffdas_calib.so"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", f"-I{ROOT / 'include'}",
          f"-I{CSRC}", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
# packed FP32 instructions are kept out of every device code object
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++",
               "-c", str(src), "-o",
               str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if verbose and warn:
        print(warn, file=sys.stderr)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link libffdas.so; always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs), "-Wl,--no-undefined", "-lamdhip64", "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_calib(verbose)
    return LIB


def build_calib(verbose: bool = False) -> Path:
    """Compile the LDS atomic-rate probe into its own shared library."""
    tmp = CALIB_LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *COMMON, *DEVICE, "-shared", "-x", "hip", str(CALIB_SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"calibration probe build failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CALIB_LIB)
    return CALIB_LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for lib in (LIB, CALIB_LIB):
        if lib.exists():
            lib.unlink()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
