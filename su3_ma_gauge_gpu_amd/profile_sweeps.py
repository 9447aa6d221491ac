"""Run the dominant kernels of the 32^4 benchmark a fixed number of times,
for kernel traces and hardware-counter runs:

* one device-to-device copy of the FP64 18-real configuration (a plain copy
  of known size, 4 * V * 18 * 8 bytes read and written), three times;
* `--sweeps` FP32 gauge-transformation sweeps on read-only compact FP32
  links, FP32 arithmetic (the mixed driver's default FP32 kernel), the
  same with FP64 arithmetic, and on the links in 24-bit fixed point;
* `--sweeps` whole sweeps of the 24-bit kernel as the mixed driver runs a
  sweep without a check: a wavefront over timeslices, with the slab the
  library's timing picks for this lattice (or the two kernels);
* `--sweeps` FP32 link sweeps with gauge accumulation on the compact FP32
  layout (the mixed driver's link mode);
* `--sweeps` fused-check FP64 sweeps on the drivers' tiled copy of the links
  (the FP64 kernel of gf_run_fast and of the link mode's FP64 phase);
* `--sweeps` fused-check FP64 gauge-transformation sweeps on the same tiled
  links held read only (the FP64 phase of the mixed run's gauge modes);
* `--sweeps` drop-in (reference-order) FP64 sweeps through gf_sweep on the
  18-real GpuPattern buffer;
* `--sweeps` launches of the drop-in quality kernel (GaugeFixingStats'
  per-site pass) on the same buffer.

Usage: python su3_ma_gauge_gpu_amd/profile_sweeps.py [--dims 32 32 32 32] [--sweeps 10] [--gauge landau|mag]
       [--block-order 0] [--slab 0] [--lag 256]
(--dims 48 48 48 48: a working set far beyond the 256 MB Infinity Cache;
--gauge mag: the same kernels with MAG's update and check).  The
gauge-transformation sweeps run in the XCD block order --block-order (0: one
pass over each XCD's timeslices, n > 0: passes of n, -1: the library's tuner
picks, timing candidates first) and the wavefront sweeps with slabs of
--slab timeslices (0: two kernels, -1: tuned, 1000000: the pipelined sweep,
one launch per sweep, with --lag extra even blocks ahead of the odd ones);
fixed settings keep the profiler's view to the kernels themselves.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))

import su3_ma_gauge_gpu_amd as gf  # noqa: E402
from su3_ma_gauge_gpu_amd import _lib  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs=4, default=[32, 32, 32, 32])
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--gauge", choices=["landau", "mag"], default="landau")
    ap.add_argument("--block-order", type=int, default=0)
    ap.add_argument("--slab", type=int, default=0)
    ap.add_argument("--lag", type=int, default=256)
    args = ap.parse_args()
    _lib.call("gf_set_block_order", args.block_order, None)
    _lib.call("gf_set_wavefront", args.slab, None)
    _lib.call("gf_set_pipelined_lag", args.lag, None)
    dims = tuple(args.dims)
    gauge = gf.MAG if args.gauge == "mag" else gf.LANDAU
    kernels = gf.MAGKernelsSU3 if args.gauge == "mag" else gf.LandauKernelsSU3
    lat64 = gf.Lattice(dims, gf.F64, gf.PATTERN18)
    lat32 = gf.Lattice(dims, gf.F32, gf.R12)
    U0 = gf.make_config(dims, 0.35 if args.gauge == "mag" else 0.25, 1, 2)
    src = gf.DeviceBuffer.from_numpy(U0)
    work = gf.DeviceBuffer(lat64.nbytes)
    for _ in range(3):
        work.copy_from(src)
    print(f"copy: {lat64.nbytes} bytes x 3", flush=True)
    u32 = gf.DeviceBuffer(lat32.nbytes)
    _lib.call("gf_convert", C.byref(lat64.struct), C.c_void_p(src.ptr), C.byref(lat32.struct), C.c_void_p(u32.ptr), 0,
              None)
    ms = C.c_float()
    for variant, what in ((2, "fp32 gauge-transformation sweep"), (3, "fp32 gauge-transformation sweep, fp64 arithmetic"),
                          (4, "fp32 gauge-transformation sweep, 24-bit fixed-point links"),
                          (6, "the same, each sweep as the mixed run's FP32 phase runs it (per half sweep)"),
                          (1, "fp32 link sweep with gauge accumulation")):
        _lib.call("gf_time_sweeps", C.byref(lat32.struct), C.c_void_p(u32.ptr), gauge, args.sweeps, variant,
                  C.byref(ms), None)
        print(f"{what}: {ms.value:.4f} ms per kernel", flush=True)
    _lib.call("gf_time_sweeps", C.byref(lat64.struct), C.c_void_p(work.ptr), gauge, args.sweeps, 0, C.byref(ms),
              None)
    print(f"fp64 fused sweep (tiled copy): {ms.value:.4f} ms per kernel", flush=True)
    _lib.call("gf_time_sweeps", C.byref(lat64.struct), C.c_void_p(work.ptr), gauge, args.sweeps, 7, C.byref(ms),
              None)
    print(f"fp64 gauge-transformation sweep (tiled copy): {ms.value:.4f} ms per kernel", flush=True)
    K = kernels(lat64)
    for _ in range(args.sweeps):
        for parity in (0, 1):
            K.orStep(lat64.V // 2 // 32, 256, work, None, parity, 1.7)
    gf.synchronize()
    print("drop-in sweeps done", flush=True)
    _lib.call("gf_time_sweeps", C.byref(lat64.struct), C.c_void_p(work.ptr), gauge, args.sweeps, 5, C.byref(ms),
              None)
    print(f"drop-in quality: {ms.value:.4f} ms per kernel", flush=True)


if __name__ == "__main__":
    main()
