"""Build the gfx950 library ``libsu3gf.so`` in-tree.

HIP sources are compiled by hipcc for gfx950 only, with packed-FP32 code
generation disabled; the host-only configuration generator is compiled by g++
with OpenMP.  Objects go to ``_build/`` next to this file, the shared library
to ``libsu3gf.so`` beside it.

Usage: ``python -m su3_ma_gauge_gpu_amd.build [--verbose]``
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsu3gf.so"
INCLUDE = PKG.parent / "include"

# (source, extra flags): the exact unit keeps every multiply and add separate
HIP_SOURCES = [
    ("su3gf_exact.hip", ["-ffp-contract=off"]),
    ("su3gf_fast.hip", []),
    ("su3gf_kernels.hip", []),
    ("su3gf_gauge.hip", []),
    ("su3gf_api.hip", []),
]
HOST_SOURCES = ["su3gf_config.cpp", "su3gf_io.cpp", "su3gf_batch.cpp"]

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    f"-I{INCLUDE}",
]
GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}",
             f"-I{INCLUDE}"]

# The application on the drop-in headers (include/culgt): host code only, so
# g++ with the HIP runtime headers; the lattice extents are compile-time
# constants as in the reference, one binary per (extent, precision).
APPS = PKG / "apps"
APP_VARIANTS = [
    ("gaugefix_x8t8_f64", ["-D_X_=8", "-D_T_=8", "-DDOUBLEPRECISION"]),
    ("gaugefix_x8t8_f32", ["-D_X_=8", "-D_T_=8"]),
    # the reference's DPUPDATES=true single-precision build (apps/Makefile:46-48)
    ("gaugefix_x8t8_f32dp", ["-D_X_=8", "-D_T_=8", "-DUSE_DP_ORUPDATE", "-DUSE_DP_MICROUPDATE", "-DUSE_DP_SAUPDATE"]),
    ("gaugefix_x32t32_f64", ["-D_X_=32", "-D_T_=32", "-DDOUBLEPRECISION"]),
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _start(cmd: list[str], verbose: bool) -> subprocess.Popen:
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:2])} ... ({res.returncode})")
    if verbose and (res.stdout or res.stderr):
        print(res.stdout + res.stderr, flush=True)


def build(verbose: bool = False) -> Path:
    """Compile every source and link ``libsu3gf.so``; returns its path."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    objs, jobs = [], []
    for src, extra in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append(_start([hipcc, *HIPCC_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)], verbose))
        objs.append(str(obj))
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append(_start(["g++", *GXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], verbose))
        objs.append(str(obj))
    failed = []
    for job in jobs:
        out, _ = job.communicate()
        if job.returncode != 0:
            failed.append(out)
        elif verbose and out:
            print(out, flush=True)
    if failed:
        sys.stderr.write("\n".join(failed))
        raise RuntimeError("compilation failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lgomp"], verbose)
    os.replace(tmp, LIB)
    build_apps(verbose)
    return LIB


def build_apps(verbose: bool = False) -> list[Path]:
    """Link the gauge-fixing application variants against ``libsu3gf.so``
    into ``apps/bin/``."""
    out = APPS / "bin"
    out.mkdir(exist_ok=True)
    jobs, paths = [], []
    for name, defs in APP_VARIANTS:
        exe = out / name
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{INCLUDE}",
               *defs, str(APPS / "gaugefix.cpp"), "-o", str(exe), f"-L{PKG}", "-lsu3gf", f"-L{ROCM / 'lib'}",
               "-lamdhip64", "-Wl,-rpath,$ORIGIN/../..", f"-Wl,-rpath,{ROCM / 'lib'}"]
        jobs.append(_start(cmd, verbose))
        paths.append(exe)
    failed = [out for job in jobs for out in [job.communicate()[0]] if job.returncode != 0]
    if failed:
        sys.stderr.write("\n".join(failed))
        raise RuntimeError("application build failed")
    return paths


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
