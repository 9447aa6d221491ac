"""Benchmark: Landau OR gauge fixing of a 32^4 lattice to theta < 1e-12 on one
MI355X, mixed precision (FP32 compact links with an accumulated gauge
transformation until the FP32 theta < 1e-12, then FP64 until the FP64 theta
< 1e-12) against plain FP64, reported as site-updates/s and time-to-theta.
The schedule that switches at theta < 1e-6 is reported beside it.

One step = one complete gauge-fixing run of the same synthetic configuration
(U = exp(i sum_a c_a lambda_a/2), c_a ~ N(0, 0.25^2), seed 1, then a Haar
gauge transformation with seed 2; omega = 1.7; checked every sweep,
reprojected every 100 sweeps).  The configuration is resident in HBM before
the timed region; each step starts from a device-to-device copy of it (the
copy is inside the timed region).

Printed: one JSON line with the metric, the roofline figure of the dominant
kernel (algorithmic bytes per launch / launch time measured here with HIP
events on the launch stream) and the CPU baseline (the FP64 oracle on this
host's cores, a bounded number of 32^4 sweeps).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--dims 32 32 32 32]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import su3_ma_gauge_gpu_amd as gf  # noqa: E402
from su3_ma_gauge_gpu_amd import _lib  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dims", type=int, nargs=4, default=[32, 32, 32, 32])
    ap.add_argument("--sigma", type=float, default=0.25)
    ap.add_argument("--omega", type=float, default=1.7)
    ap.add_argument("--precision", type=float, default=1e-12)
    ap.add_argument("--switch", type=float, default=1e-12)
    ap.add_argument("--switch-northstar", type=float, default=1e-6)
    ap.add_argument("--refresh", type=int, default=50)
    ap.add_argument("--check32", type=int, default=10)
    ap.add_argument("--reproject", type=int, default=100)
    ap.add_argument("--maxiter", type=int, default=20000)
    ap.add_argument("--kernel-sweeps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    return ap.parse_args()


class Events:
    def __init__(self):
        self.a, self.b = C.c_void_p(), C.c_void_p()
        _lib.call("gf_event_create", C.byref(self.a))
        _lib.call("gf_event_create", C.byref(self.b))

    def start(self, stream):
        _lib.call("gf_event_record", self.a, stream)

    def stop_ms(self, stream) -> float:
        _lib.call("gf_event_record", self.b, stream)
        ms = C.c_float()
        _lib.call("gf_event_elapsed_ms", self.a, self.b, C.byref(ms))
        return float(ms.value)


def kernel_time_ms(lat: gf.Lattice, U: gf.DeviceBuffer, stream, nsweeps: int, with_gauge: bool) -> float:
    """Average duration of one fused-check parity sweep kernel (the kernel the
    drivers run), from HIP events around nsweeps back-to-back sweeps
    (2 * nsweeps launches) on `stream`."""
    ms = C.c_float()
    _lib.call("gf_time_sweeps", C.byref(lat.struct), C.c_void_p(U.ptr), gf.LANDAU, int(nsweeps), int(with_gauge),
              C.byref(ms), stream)
    return float(ms.value)


def dropin_kernel_ms(lat: gf.Lattice, U: gf.DeviceBuffer, stream, nsweeps: int) -> float:
    """Average duration of one drop-in (reference-order) parity sweep kernel,
    gf_sweep, from HIP events around nsweeps back-to-back sweeps on `stream`."""
    ev = Events()
    for parity in (0, 1):
        _lib.call("gf_sweep", C.byref(lat.struct), C.c_void_p(U.ptr), C.c_void_p(U.ptr), parity, gf.LANDAU, gf.OR,
                  C.c_float(1.7), 0, 0, 32, stream)
    ev.start(stream)
    for _ in range(nsweeps):
        for parity in (0, 1):
            _lib.call("gf_sweep", C.byref(lat.struct), C.c_void_p(U.ptr), C.c_void_p(U.ptr), parity, gf.LANDAU,
                      gf.OR, C.c_float(1.7), 0, 0, 32, stream)
    return ev.stop_ms(stream) / (2 * nsweeps)


def cpu_baseline(dims, U_host: np.ndarray, seconds: float) -> dict:
    """The FP64 oracle (the reference sweep restated in C, OpenMP over the sites
    of a parity) on this host's cores, on a bounded number of sweeps."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import LANDAU, OR, Oracle

    oracle = Oracle("f64")
    cores = len(os.sched_getaffinity(0))
    oracle.set_threads(cores)
    U = U_host.copy()
    V = int(np.prod(dims))
    n = 0
    t0 = time.perf_counter()
    while True:
        oracle.sweep(LANDAU, OR, dims, U, 0, 1.7)
        oracle.sweep(LANDAU, OR, dims, U, 1, 1.7)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 1000:
            break
    return {
        "value": n * V / el,
        "unit": "site-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} FP64 OR sweeps of the {'x'.join(map(str, dims))} config ({el:.1f} s), "
                  f"oracle with {cores} OpenMP threads",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build is single-GPU: "
              "see DESIGN.md for how the sweep would shard)", file=sys.stderr)
        sys.exit(2)
    dims = tuple(args.dims)
    V = int(np.prod(dims))
    print(f"bench: generating {dims} config", file=sys.stderr, flush=True)
    U0 = gf.make_config(dims, args.sigma, 1, 2)
    lat64 = gf.Lattice(dims, gf.F64, gf.PATTERN18)
    lat32 = gf.Lattice(dims, gf.F32, gf.R12)
    src = gf.DeviceBuffer.from_numpy(U0)
    work = gf.DeviceBuffer(lat64.nbytes)
    stream = C.c_void_p()
    _lib.call("gf_stream_create", C.byref(stream))

    def reset():
        # stream-ordered copy of the resident input: part of the timed step
        _lib.call("gf_memcpy_async", C.c_void_p(work.ptr), C.c_void_p(src.ptr), C.c_size_t(lat64.nbytes), stream)

    def one_mixed():
        reset()
        return gf.run_mixed(lat64, work, gf.LANDAU, args.omega, args.switch, args.precision, args.maxiter,
                            args.refresh, args.check32, args.reproject, stream=stream)

    for w in range(args.warmup):
        r = one_mixed()
        print(f"bench: warmup {w}: {r}", file=sys.stderr, flush=True)
    _lib.call("gf_synchronize")
    ev = Events()
    results = []
    ev.start(stream)
    t0 = time.perf_counter()
    for s in range(args.steps):
        results.append(one_mixed())
    ms_total = ev.stop_ms(stream)
    wall = time.perf_counter() - t0
    _lib.call("gf_synchronize")
    r = results[-1]
    sweeps = sum(x.sweeps for x in results)
    su_per_s = sweeps * V / (ms_total / 1e3)
    print(f"bench: mixed {r}", file=sys.stderr, flush=True)

    # plain FP64 reference run (same driver, FP64 18-real layout throughout)
    fp64 = None
    if not args.no_fp64:
        reset()
        fp64 = gf.run_fast(lat64, work, gf.LANDAU, args.omega, args.precision, args.maxiter, args.reproject,
                           stream=stream)
        print(f"bench: fp64 {fp64}", file=sys.stderr, flush=True)

    # the same run with the FP32 -> FP64 switch at theta < 1e-6
    ns = None
    if not args.no_fp64:
        reset()
        ns = gf.run_mixed(lat64, work, gf.LANDAU, args.omega, args.switch_northstar, args.precision, args.maxiter,
                          args.refresh, args.check32, args.reproject, stream=stream)
        print(f"bench: mixed, switch {args.switch_northstar:g}: {ns}", file=sys.stderr, flush=True)

    # kernel timing for the roofline (both precisions, on the launch stream)
    U32 = gf.DeviceBuffer(lat32.nbytes)
    _lib.call("gf_convert", C.byref(lat64.struct), C.c_void_p(src.ptr), C.byref(lat32.struct), C.c_void_p(U32.ptr), 0,
              stream)
    reset()
    k32 = kernel_time_ms(lat32, U32, stream, args.kernel_sweeps, True)
    k64 = kernel_time_ms(lat64, work, stream, args.kernel_sweeps, False)
    reset()
    kdrop = dropin_kernel_ms(lat64, work, stream, args.kernel_sweeps)
    # algorithmic bytes per parity launch: V/2 site-updates x 16 R sizeof(real), R = 12;
    # the FP32 kernel also reads and writes the 12-float gauge transformation of its V/2 sites
    bytes32 = (V // 2) * (16 * 12 * 4 + 2 * 12 * 4)
    bytes64 = (V // 2) * 16 * 12 * 8
    t32 = r.ms_fp32 / max(r.sweeps_fp32, 1)
    t64 = (r.ms_total - r.ms_fp32) / max(r.sweeps - r.sweeps_fp32, 1)
    dom32 = r.sweeps_fp32 * k32 >= (r.sweeps - r.sweeps_fp32) * k64
    kname, kms, kbytes = ("sweep_f32_r12_gauge", k32, bytes32) if dom32 else ("sweep_f64_pattern18", k64, bytes64)
    achieved = kbytes / (kms / 1e3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(kname, {}).get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dims, U0, args.cpu_seconds)

    line = {
        "metric": "Landau OR site-updates/s + time to theta<1e-12 at 32^4: mixed vs FP64",
        "value": su_per_s,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"landau-or-{'x'.join(map(str, dims))}-mixed",
            "dims": list(dims),
            "sigma": args.sigma,
            "seed": 1,
            "haar_seed": 2,
            "omega": args.omega,
            "precision": args.precision,
            "switch_precision": args.switch,
            "check32": args.check32,
            "checkprecision": 1,
            "refresh": args.refresh,
            "reproject": args.reproject,
        },
        "time_to_theta_mixed_s": r.ms_total / 1e3,
        "sweeps_mixed": r.sweeps,
        "sweeps_fp32": r.sweeps_fp32,
        "gff_mixed": r.gff,
        "theta_mixed": r.theta,
        "time_to_theta_fp64_s": None if fp64 is None else fp64.ms_total / 1e3,
        "sweeps_fp64": None if fp64 is None else fp64.sweeps,
        "gff_fp64": None if fp64 is None else fp64.gff,
        "gff_rel_diff_mixed_vs_fp64": None if fp64 is None else abs(r.gff - fp64.gff) / abs(fp64.gff),
        "speedup_mixed_vs_fp64": None if fp64 is None else fp64.ms_total / r.ms_total,
        "mixed_switch_1e-6": None if ns is None else {
            "time_to_theta_s": ns.ms_total / 1e3, "sweeps": ns.sweeps, "sweeps_fp32": ns.sweeps_fp32,
            "gff": ns.gff, "speedup_vs_fp64": fp64.ms_total / ns.ms_total},
        "site_updates_per_s_fp64": None if fp64 is None else fp64.sweeps * V / (fp64.ms_total / 1e3),
        "sweep_ms_in_run": {"fp32": t32, "fp64": t64},
        "kernel_ms": {"sweep_f32_r12_gauge": k32, "sweep_f64_pattern18": k64, "dropin_sweep_f64_pattern18": kdrop},
        "roofline_dropin_f64_frac": bytes64 / (kdrop / 1e3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": kbytes,
        },
        "roofline_fp64_kernel_frac": bytes64 / (k64 / 1e3) / 1e9 / HBM_PEAK_GBS,
        "roofline_fp32_kernel_frac": bytes32 / (k32 / 1e3) / 1e9 / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
        "wall_s": wall,
    }
    _lib.call("gf_stream_destroy", stream)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
