"""Benchmark: Landau OR gauge fixing of a 32^4 lattice to theta < 1e-12 on one
MI355X, mixed precision (an FP32 phase that sweeps the gauge transformation
g(x) on a read-only FP32 copy of the links until its theta < 1e-12 or stops
improving; then g applied to the FP64 links and FP64 sweeps until the FP64
theta < 1e-12) against plain FP64, reported as site-updates/s and
time-to-theta.  The other single-precision modes and the schedule that
switches at theta < 1e-6 are reported beside it.

One step = one complete gauge-fixing run of the same synthetic configuration
(U = exp(i sum_a c_a lambda_a/2), c_a ~ N(0, 0.25^2), seed 1, then a Haar
gauge transformation with seed 2; omega = 1.7; checked every sweep,
reprojected every 100 sweeps).  The configuration is resident in HBM before
the timed region; each step starts from a device-to-device copy of it (the
copy is inside the timed region).

Printed: one JSON line with the metric, the roofline figure of the dominant
kernel (algorithmic bytes per launch / launch time measured here with HIP
events on the launch stream) and the CPU baseline (the FP64 oracle on this
host's cores, a bounded number of 32^4 sweeps).  Under "workloads" the same
line carries one run of each other workload, outside the timed steps:
Coulomb 32^3 x 64 (every slice at once), the gauge copies (4 each of 16^4 to
32^4, annealing then OR), MAG to theta < 1e-12 at 24^4 then U1xU1 (with MAG's
kernels timed at 48^4) and Landau at 48^4, where the lattice streams from
HBM instead of the Infinity Cache (--no-extras leaves them out).

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
MODE_NAMES = {0: "links_f32+g", 1: "gauge_f32", 2: "gauge_f32_f64arith", 3: "gauge_q24", 4: "fp64_gauge_sweep"}


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
    ap.add_argument("--early-switch", type=float, default=1e-6)
    ap.add_argument("--refresh", type=int, default=50)
    ap.add_argument("--check32", type=int, default=5)
    ap.add_argument("--reproject", type=int, default=100)
    ap.add_argument("--maxiter", type=int, default=20000)
    ap.add_argument("--kernel-sweeps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="leave out the other workloads (Coulomb, gauge copies, MAG, Landau 48^4) of the default line")
    ap.add_argument("--mode", type=int, default=3, help="0 links+g, 1 g (FP32 arithmetic), 2 g (FP64 arithmetic), 3 g on 24-bit fixed-point links")
    ap.add_argument("--copies", type=int, default=4)
    ap.add_argument("--sa-steps", type=int, default=3000)
    ap.add_argument("--sa-mode", type=int, default=2,
                    help="0 FP64 link sweeps, 1 FP32 gauge-transformation sweeps, 2 the same on 24-bit links")
    ap.add_argument("--mag-precision", type=float, default=1e-12)
    ap.add_argument("--maxiter-mag", type=int, default=600000)
    ap.add_argument("--u1xu1-sweeps", type=int, default=1000)
    ap.add_argument("--mag-runs", default="both", choices=["both", "mixed", "fp64", "fp64-gauge"],
                    help="--workload mag: the mixed and the all-FP64 run, or one of them (each takes minutes at 48^4); "
                         "fp64-gauge: all-FP64 on the FP64 gauge-transformation sweep")
    ap.add_argument("--workload", default="landau",
                    choices=["landau", "coulomb", "gribov", "mag", "small-copies"],
                    help="landau: the 32^4 headline run (default); coulomb: 32^3x64 all slices at once; "
                         "gribov: SA + OR gauge copies of 16^4..32^4; mag: 48^4 MAG then U1xU1; "
                         "small-copies: annealing of 16 copies each of 8^4 and 12^4, a launch per copy "
                         "against one launch for all")
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


def sweep_path(slab: int) -> str:
    """How the library runs a gauge-transformation sweep without a check, from
    its wavefront choice."""
    if slab == gf.WAVEFRONT_PIPELINED:
        return "pipelined"
    return f"wavefront, slabs of {slab} timeslices" if slab > 0 else "two kernels"


def pipelined_launches() -> int:
    n = C.c_longlong()
    _lib.call("gf_pipelined_launches", C.byref(n))
    return n.value


def kernel_time_ms(lat: gf.Lattice, U: gf.DeviceBuffer, stream, nsweeps: int, variant: int,
                   gauge: int = gf.LANDAU) -> float:
    """Average duration of one parity sweep kernel as the drivers run it, from
    HIP events around nsweeps back-to-back sweeps (2 * nsweeps launches) on
    `stream`.  variant: 0 link sweep in the precision of `lat` with the fused
    check, 1 FP32 links + g, 2 / 3 gauge-transformation sweep in FP32 / FP64
    arithmetic, 4 the FP32 gauge-transformation sweep on 24-bit links, 5 the
    drop-in quality kernel (nsweeps launches; the time is per launch)."""
    ms = C.c_float()
    _lib.call("gf_time_sweeps", C.byref(lat.struct), C.c_void_p(U.ptr), int(gauge), int(nsweeps), int(variant),
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


def cpu_baseline(dims, U_host: np.ndarray, seconds: float, threads: int | None = None) -> dict:
    """The FP64 oracle (the reference sweep restated in C, OpenMP over the sites
    of a parity) on this host's cores (or `threads` of them), on a bounded
    number of sweeps."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import LANDAU, OR, Oracle

    oracle = Oracle("f64")
    cores = threads or len(os.sched_getaffinity(0))
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


def _stream():
    stream = C.c_void_p()
    _lib.call("gf_stream_create", C.byref(stream))
    return stream


def pretune(lat: gf.Lattice, src: gf.DeviceBuffer, gauge: int, modes, stream, sa_mode=None) -> float:
    """The library's one-time tuning for a kind of lattice, outside the timed
    runs: the first mixed run of a (shape, gauge, link form) times candidate
    sweep paths for its FP32 phase, its FP64-arithmetic stage and its FP64
    phase (about 0.1-0.3 s at 32^4), and the first annealing schedule those of
    its sweeps.  A two-sweep run_mixed per mode in `modes` (and a two-step
    schedule in `sa_mode`) on a scratch copy of `src` does all of that.
    Returns the seconds it took, which the lines report as `tuning_s`."""
    scratch = gf.DeviceBuffer(lat.nbytes)
    _lib.call("gf_synchronize")
    t0 = time.perf_counter()
    for m in modes:
        _lib.call("gf_memcpy_async", C.c_void_p(scratch.ptr), C.c_void_p(src.ptr), C.c_size_t(lat.nbytes), stream)
        gf.run_mixed(lat, scratch, gauge, 1.7, 1e-12, 1e-12, 2, mode=m, stream=stream)
    if sa_mode is not None:
        _lib.call("gf_memcpy_async", C.c_void_p(scratch.ptr), C.c_void_p(src.ptr), C.c_size_t(lat.nbytes), stream)
        gf.run_sa(lat, scratch, gauge, 1.45, 0.01, 2, 3, 100, mode=sa_mode, stream=stream)
    _lib.call("gf_synchronize")
    return time.perf_counter() - t0


def _base_line(metric, value, unit, ms_per_step, steps, warmup, config, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic", "config": config}


def slice_quality(lat: gf.Lattice, U: gf.DeviceBuffer, stream):
    """Per-timeslice Coulomb theta and gff of a 4D lattice (prefactors of
    CoulombKernelsSU3.hxx:71-78 per slice): per-site values from
    gf_quality_per_site, summed per slice on the host."""
    V = lat.V
    dG = gf.DeviceBuffer(8 * V)
    dA = gf.DeviceBuffer(8 * V)
    _lib.call("gf_quality_per_site", C.byref(lat.struct), gf.COULOMB, C.c_void_p(U.ptr), C.c_void_p(dG.ptr),
              C.c_void_p(dA.ptr), stream)
    _lib.call("gf_stream_synchronize", stream)
    g = dG.download(np.float64, V).reshape(2, lat.dims[0], -1).sum(axis=(0, 2))
    a = dA.download(np.float64, V).reshape(2, lat.dims[0], -1).sum(axis=(0, 2))
    V3 = V // lat.dims[0]
    return a / (3 * V3), g / (9 * V3)


def workload_coulomb(args):
    """Coulomb gauge of a 32^3 x 64 lattice (sigma 0.3, seed 1, Haar seed 2),
    every timeslice at once: mixed (FP32 gauge-transformation sweeps, then
    FP64) against all-FP64, each slice to theta < 1e-12."""
    print(json.dumps(coulomb_line(args)), flush=True)


def coulomb_line(args) -> dict:
    dims = (64, 32, 32, 32)
    V = int(np.prod(dims))
    U0 = gf.make_config(dims, 0.3, 1, 2)
    lat = gf.Lattice(dims)
    src = gf.DeviceBuffer.from_numpy(U0)
    work = gf.DeviceBuffer(lat.nbytes)
    stream = _stream()

    def reset():
        _lib.call("gf_memcpy_async", C.c_void_p(work.ptr), C.c_void_p(src.ptr), C.c_size_t(lat.nbytes), stream)

    def mixed():
        reset()
        return gf.run_mixed(lat, work, gf.COULOMB, args.omega, args.switch, args.precision, args.maxiter,
                            mode=args.mode, check32=args.check32, stream=stream)

    # the library picks its wavefront slab for a lattice on its first run
    reset()
    gf.run_mixed(lat, work, gf.COULOMB, args.omega, args.switch, args.precision, 2, mode=args.mode, stream=stream)
    for _ in range(args.warmup):
        mixed()
    n_pipe = pipelined_launches()
    rs = [mixed() for _ in range(args.steps)]
    n_pipe = (pipelined_launches() - n_pipe) // max(1, args.steps)
    r = rs[-1]
    th_m, g_m = slice_quality(lat, work, stream)
    slab = C.c_int()
    _lib.call("gf_wavefront_choice", C.byref(lat.struct), gf.COULOMB, args.mode, C.byref(slab))
    # the same run with every sweep as two kernels
    prev = C.c_int()
    _lib.call("gf_set_wavefront", 0, C.byref(prev))
    r2 = mixed()
    _lib.call("gf_set_wavefront", prev.value, None)
    # the workload's stated schedule: FP32 12-real links until theta < 1e-6,
    # then FP64 sweeps (on the gauge transformation); its mode's first run
    # would tune its sweep paths inside its timing
    tuning_s = pretune(lat, src, gf.COULOMB, [gf.MIXED_GAUGE32], stream) if args.mode != gf.MIXED_GAUGE32 else 0.0
    reset()
    st = gf.run_mixed(lat, work, gf.COULOMB, args.omega, args.early_switch, args.precision, args.maxiter,
                      mode=gf.MIXED_GAUGE32, check32=args.check32, stream=stream)
    th_s, g_s = slice_quality(lat, work, stream)
    reset()
    f = gf.run_fast(lat, work, gf.COULOMB, args.omega, args.precision, args.maxiter, args.reproject, stream=stream)
    th_f, g_f = slice_quality(lat, work, stream)
    ms = sum(x.ms_total for x in rs) / len(rs)
    line = _base_line("Coulomb OR site-updates/s + time to theta<1e-12 per slice at 32^3x64: mixed vs FP64",
                      r.sweeps * V / (ms / 1e3), "site-updates/s", ms, args.steps, args.warmup,
                      {"workload": "coulomb-or-32x32x32-x64-slices-mixed", "dims": list(dims), "sigma": 0.3,
                       "seed": 1, "haar_seed": 2, "omega": args.omega, "precision": args.precision,
                       "switch_precision": args.switch, "mode": MODE_NAMES[args.mode]})
    line.update({
        "time_to_theta_mixed_s": ms / 1e3, "sweeps_mixed": r.sweeps, "sweeps_fp32": r.sweeps_fp32,
        "sweep_path": sweep_path(slab.value), "pipelined_sweeps_in_mixed_run": n_pipe,
        "time_to_theta_mixed_two_kernel_sweeps_s": r2.ms_total / 1e3,
        "mixed_identical_with_and_without_wavefront": (r.sweeps, r.gff, r.theta) == (r2.sweeps, r2.gff, r2.theta),
        "time_to_theta_fp64_s": f.ms_total / 1e3, "sweeps_fp64": f.sweeps,
        "speedup_mixed_vs_fp64": f.ms_total / ms,
        "max_slice_theta_mixed": float(th_m.max()), "max_slice_theta_fp64": float(th_f.max()),
        "max_slice_gff_rel_diff": float(np.max(np.abs(g_m - g_f) / np.abs(g_f))),
        "gff_mixed": r.gff, "gff_fp64": f.gff,
        "mixed_fp32_links_switch_1e-6": {
            "time_to_theta_s": st.ms_total / 1e3, "sweeps": st.sweeps, "sweeps_fp32": st.sweeps_fp32,
            "speedup_vs_fp64": f.ms_total / st.ms_total, "max_slice_theta": float(th_s.max()),
            "max_slice_gff_rel_diff": float(np.max(np.abs(g_s - g_f) / np.abs(g_f))), "tuning_s": tuning_s},
    })
    return line


def workload_mag(args):
    """MAG of a 48^4 lattice (sigma 0.35, seed 1, Haar seed 2), the full
    schedule: mixed (FP32 phase in --mode, switch at theta < 1e-6, then FP64)
    against all-FP64 to theta < --mag-precision, then U1xU1 OR sweeps on the
    MAG result; MAG's sweep kernels timed one by one.  At this size OR needs
    a few hundred thousand sweeps: expect minutes per run."""
    line = mag_line(args, 48, args.maxiter_mag, args.mag_runs)
    if args.mag_runs == "both":
        line.update(mag_kernels(args, 48))
    print(json.dumps(line), flush=True)


def _run_chunked(lat, work, gauge, args, prec, maxiter, stream, what, gauge_sweep=False):
    """FP64 OR to theta < prec in chunks of 20000 sweeps (a progress line
    after each): gf_run_fast (link sweeps), or with gauge_sweep the FP64
    gauge-transformation sweep (gf_run_mixed without its FP32 phase).  One
    run, as the sweeps do not depend on where a chunk ends but for a
    reprojection at each chunk boundary.  Returns (sweeps, ms, last result)."""
    sweeps, ms, r = 0, 0.0, None
    while sweeps < maxiter:
        n = min(20000, maxiter - sweeps)
        if gauge_sweep:
            r = gf.run_mixed(lat, work, gauge, args.omega, prec, prec, n, mode=gf.MIXED_FP64,
                             reproject=args.reproject, stream=stream)
        else:
            r = gf.run_fast(lat, work, gauge, args.omega, prec, n, args.reproject, stream=stream)
        sweeps += r.sweeps
        ms += r.ms_total
        print(f"bench: {what}: {sweeps} sweeps, theta {r.theta:.3e}, {ms / 1e3:.1f} s", file=sys.stderr, flush=True)
        if r.converged:
            break
    return sweeps, ms, r


def mag_quality(lat, U, stream):
    """MAG gff and theta of a lattice (GaugeFixingStats with MAGKernelsSU3's
    prefactors, the drop-in quality kernel)."""
    st = gf.GaugeFixingStats(gf.MAGKernelsSU3, U, lat, stream=stream)
    st.generateGaugeQuality()
    g, t = st.getCurrentGff(), st.getCurrentA()
    st.close()
    return g, t


def mag_line(args, L: int, maxiter: int, runs: str = "both") -> dict:
    """MAG of an L^4 lattice (sigma 0.35, seed 1, Haar seed 2) to theta <
    --mag-precision: mixed (the FP32 phase in --mode until theta < 1e-6, then
    FP64 sweeps) against all-FP64 (runs: "both", or one of "mixed", "fp64",
    for runs too long to make together); then --u1xu1-sweeps U1xU1 OR sweeps
    on each MAG result, which leave MAG's functional and theta unchanged (a
    diagonal gauge rotation changes neither |U_jj| nor |X_k|)."""
    dims = (L, L, L, L)
    V = int(np.prod(dims))
    U0 = gf.make_config(dims, 0.35, 1, 2)
    lat = gf.Lattice(dims)
    src = gf.DeviceBuffer.from_numpy(U0)
    work = gf.DeviceBuffer(lat.nbytes)
    stream = _stream()
    prec = args.mag_precision

    def reset():
        _lib.call("gf_memcpy_async", C.c_void_p(work.ptr), C.c_void_p(src.ptr), C.c_size_t(lat.nbytes), stream)

    mixed = fp64 = m = f = None
    fixed = []  # (name, buffer holding that run's MAG result)
    # the first mixed run of this lattice would tune its sweep paths (the FP32
    # phase's, its FP64-arithmetic stage's and the FP64 phase's) inside its timing
    tuning_s = pretune(lat, src, gf.MAG, [args.mode], stream) if runs in ("both", "mixed") else 0.0
    if runs in ("both", "mixed"):
        # the FP32 phase to 1e-6 (gf_run_mixed with precision 1e-6 ends on its
        # first FP64 check), then FP64 sweeps to the precision as the mixed
        # driver's FP64 phase runs them (the FP64 gauge-transformation sweep
        # in the gauge modes), in chunks
        reset()
        m32 = gf.run_mixed(lat, work, gf.MAG, args.omega, 1e-6, 1e-6, maxiter, mode=args.mode, refresh=args.refresh,
                           check32=args.check32, reproject=args.reproject, stream=stream)
        print(f"bench: mag {L}^4 mixed, FP32 phase: {m32}", file=sys.stderr, flush=True)
        n64, ms64, m = _run_chunked(lat, work, gf.MAG, args, prec, maxiter - m32.sweeps, stream,
                                    f"mag {L}^4 mixed, FP64", gauge_sweep=args.mode != 0)
        mixed = {"time_to_theta_s": (m32.ms_total + ms64) / 1e3, "sweeps": m32.sweeps + n64,
                 "sweeps_fp32": m32.sweeps_fp32, "time_fp32_s": m32.ms_fp32 / 1e3, "converged": m.converged,
                 "gff": m.gff, "theta": m.theta}
        if runs == "both":
            keep = gf.DeviceBuffer(lat.nbytes)
            _lib.call("gf_memcpy_async", C.c_void_p(keep.ptr), C.c_void_p(work.ptr), C.c_size_t(lat.nbytes), stream)
            fixed.append(("mixed", keep))
        else:
            fixed.append(("mixed", work))
    if runs in ("both", "fp64"):
        reset()
        nf, msf, f = _run_chunked(lat, work, gf.MAG, args, prec, maxiter, stream, f"mag {L}^4 FP64")
        fp64 = {"time_to_theta_s": msf / 1e3, "sweeps": nf, "converged": f.converged, "gff": f.gff, "theta": f.theta}
        fixed.insert(0, ("fp64", work))
    if runs == "fp64-gauge":
        # all-FP64 on the FP64 gauge-transformation sweep (gf_run_mixed
        # GF_MIXED_FP64: the mixed run's FP64 phase with its check rules)
        tuning_s = pretune(lat, src, gf.MAG, [gf.MIXED_FP64], stream)
        reset()
        nf, msf, f = _run_chunked(lat, work, gf.MAG, args, prec, maxiter, stream, f"mag {L}^4 FP64 gauge sweep",
                                  gauge_sweep=True)
        fp64 = {"time_to_theta_s": msf / 1e3, "sweeps": nf, "converged": f.converged, "gff": f.gff, "theta": f.theta,
                "driver": "gf_run_mixed GF_MIXED_FP64"}
        fixed.insert(0, ("fp64", work))
    # U1xU1 on each MAG result
    u1 = {}
    for name, buf in fixed:
        g0, t0 = mag_quality(lat, buf, stream)
        if name == "fp64":
            u = gf.run_fast(lat, buf, gf.U1XU1, args.omega, 1e-30, args.u1xu1_sweeps, args.reproject, stream=stream)
        else:
            u = gf.run_mixed(lat, buf, gf.U1XU1, args.omega, 1e-30, 1e-30, args.u1xu1_sweeps, mode=args.mode,
                             refresh=args.refresh, check32=args.check32, reproject=args.reproject, stream=stream)
        g1, t1 = mag_quality(lat, buf, stream)
        u1[name] = {"sweeps": u.sweeps, "time_s": u.ms_total / 1e3, "u1xu1_gff": u.gff,
                    "mag_gff_rel_change": abs(g1 - g0) / g0, "mag_theta_before": t0, "mag_theta_after": t1}
        print(f"bench: u1xu1 on the {name} MAG result ({L}^4): {u}", file=sys.stderr, flush=True)
    del fixed
    main_run = mixed if mixed is not None else fp64
    line = _base_line(f"MAG OR time to theta<{prec:g} at {L}^4: mixed vs FP64, then U1xU1",
                      main_run["sweeps"] * V / main_run["time_to_theta_s"], "site-updates/s",
                      main_run["time_to_theta_s"] * 1e3, 1, 0,
                      {"workload": f"mag-or-{L}x{L}x{L}x{L}-mixed+u1xu1", "dims": list(dims), "sigma": 0.35,
                       "seed": 1, "haar_seed": 2, "omega": args.omega, "precision": prec, "switch_precision": 1e-6,
                       "mode": MODE_NAMES[args.mode], "u1xu1_sweeps": args.u1xu1_sweeps, "runs": runs})
    line.update({"mixed": mixed, "fp64": fp64, "u1xu1": u1, "tuning_s": tuning_s})
    if mixed is not None and fp64 is not None:
        both = mixed["converged"] and fp64["converged"]
        line.update({
            "gff_rel_diff_mixed_vs_fp64": abs(m.gff - f.gff) / abs(f.gff),
            # a time-to-theta ratio only when both runs reached the precision
            "speedup_mixed_vs_fp64": fp64["time_to_theta_s"] / mixed["time_to_theta_s"] if both else None,
            "u1xu1_gff_rel_diff_mixed_vs_fp64": abs(u1["mixed"]["u1xu1_gff"] - u1["fp64"]["u1xu1_gff"])
            / abs(u1["fp64"]["u1xu1_gff"]),
        })
    return line


def mag_kernels(args, L: int) -> dict:
    """MAG's sweep kernels at L^4, one parity launch each (V/2 site-updates),
    and their algorithmic bytes as a fraction of 8 TB/s (bytes per
    site-update as in the Landau kernel table: the update rule does not change
    what is read)."""
    dims = (L, L, L, L)
    V = int(np.prod(dims))
    lat = gf.Lattice(dims)
    src = gf.DeviceBuffer.from_numpy(gf.make_config(dims, 0.35, 1, 2))
    stream = _stream()
    lat32 = gf.Lattice(dims, gf.F32, gf.R12)
    U32 = gf.DeviceBuffer(lat32.nbytes)
    kms, frac = {}, {}
    for var, name, bsu in ((4, "gauge_sweep_q24", 432), (2, "gauge_sweep_f32", 528), (1, "sweep_f32_r12_gauge", 864),
                           (0, "sweep_f64_tiled12", 1536), (7, "gauge_sweep_f64_tiled12", 1056)):
        if var in (1, 2, 4):
            _lib.call("gf_convert", C.byref(lat.struct), C.c_void_p(src.ptr), C.byref(lat32.struct),
                      C.c_void_p(U32.ptr), 0, stream)
            ms = kernel_time_ms(lat32, U32, stream, args.kernel_sweeps, var, gf.MAG)
        else:
            ms = kernel_time_ms(lat, src, stream, args.kernel_sweeps, var, gf.MAG)
        kms[name] = ms
        frac[name] = (V / 2) * bsu / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    return {f"kernel_ms_{L}": kms, f"kernel_hbm_frac_{L}": frac}


def workload_gribov(args):
    streams = [_stream() for _ in range(args.copies)]
    line = gribov_line(args, streams)
    line["with_sa_graph"] = gribov_line(args, streams, graph=True)
    for key, kw in (("one_launch", {}), ("one_launch_sa_graph", {"graph": True}),
                    ("one_launch_all_sixteen", {"limit": 0})):
        line[key] = gribov_one_launch_line(args, streams, **kw)
        line[key]["same_best_gff_as_per_copy_launches"] = same_copies(line[key], line)
    line["small_copies_annealing"] = {f"{L}^4": small_copies_line(args, streams, L) for L in (8, 12)}
    fp64 = gribov_fp64(args, streams[0])
    line["fp64_reference_algorithm"] = fp64
    line["speedup_vs_fp64"] = fp64["total_s"] / line["total_s"]
    print(json.dumps(line), flush=True)


def workload_small_copies(args):
    streams = [_stream() for _ in range(4)]
    out = {f"{L}^4": small_copies_line(args, streams, L) for L in (8, 12)}
    print(json.dumps({"workload": "small-copies-annealing", "sa_steps": args.sa_steps, "sa_mode": args.sa_mode,
                      "results": out}), flush=True)


def gribov_fp64(args, stream) -> dict:
    """The gauge copies of gribov_line as the reference apps run them, in FP64
    (the annealing on FP64 link sweeps, run_sa mode 0, then the FP64 OR of
    gf_run_fast), one copy after the other on one stream, each size after an
    untimed first run of its kind; wall time from the first sweep to the last
    copy's result, inputs resident."""
    sizes = [16, 20, 24, 32]
    total_s, per_size, haar = 0.0, {}, 10
    for L in sizes:
        dims = (L, L, L, L)
        lat = gf.Lattice(dims)
        hs = list(range(haar, haar + args.copies))
        haar += args.copies
        bufs = [gf.DeviceBuffer.from_numpy(gf.make_config(dims, 0.3, 1, h)) for h in hs]
        scratch = gf.DeviceBuffer(lat.nbytes)
        _lib.call("gf_memcpy_async", C.c_void_p(scratch.ptr), C.c_void_p(bufs[0].ptr), C.c_size_t(lat.nbytes), stream)
        gf.run_sa(lat, scratch, gf.LANDAU, 1.45, 0.01, 2, 3, args.reproject, mode=0, stream=stream)
        gf.run_fast(lat, scratch, gf.LANDAU, args.omega, args.precision, 2, args.reproject, stream=stream)
        _lib.call("gf_synchronize")
        t0 = time.perf_counter()
        gffs, sweeps = [], []
        for h, buf in zip(hs, bufs):
            gf.run_sa(lat, buf, gf.LANDAU, 1.45, 0.01, args.sa_steps, 3, args.reproject, seed=1,
                      counter=(h - 10) * 100000, mode=0, stream=stream)
            r = gf.run_fast(lat, buf, gf.LANDAU, args.omega, args.precision, args.maxiter, args.reproject,
                            stream=stream)
            gffs.append(r.gff)
            sweeps.append(r.sweeps)
        _lib.call("gf_synchronize")
        size_s = time.perf_counter() - t0
        total_s += size_s
        per_size[f"{L}^4"] = {"best_gff": max(gffs), "wall_s": size_s, "or_sweeps": sweeps}
        print(f"bench: fp64 copies {L}^4: {size_s:.2f} s, best gff {max(gffs)!r}", file=sys.stderr, flush=True)
    return {"total_s": total_s, "per_size": per_size}


def gribov_line(args, streams=None, graph=False) -> dict:
    """Gauge copies: for each of 16^4, 20^4, 24^4, 32^4 (sigma 0.3, seed 1),
    four copies with Haar seeds 10.. (one per copy); each copy is annealed
    (T 1.45 -> 0.01 in args.sa_steps steps, 3 microcanonical steps per
    heatbath step) and then OR-fixed to theta < 1e-12 (mixed precision); the
    best functional per size is kept.  The copies of a size go through the
    library's batch calls (gf_batch_run_sa, then gf_batch_run_mixed): copies
    too small to fill the GPU alone run concurrently, each on a stream and
    host thread of its own; each copy has its own RNG counter range.  The
    streams are created here, before the first device allocation of the
    workload (streams created after it have run the concurrent copies up to
    35% slower with the runtime's 4 hardware queues per process).  graph:
    each copy's annealing schedule is captured into one HIP graph and
    replayed (GF_SA_GRAPH) instead of launched kernel by kernel."""
    if streams is None:
        streams = [_stream() for _ in range(args.copies)]
    sizes = [16, 20, 24, 32]
    total_su, total_s, per_size = 0, 0.0, {}
    tuning_s = 0.0
    haar = 10
    for L in sizes:
        dims = (L, L, L, L)
        V = L ** 4
        lats = [gf.Lattice(dims)] * args.copies
        hs = list(range(haar, haar + args.copies))
        haar += args.copies
        bufs = [gf.DeviceBuffer.from_numpy(gf.make_config(dims, 0.3, 1, h)) for h in hs]
        # the first annealing schedule and mixed run of a size tune their sweep
        # paths; done here, on a scratch copy, outside the timed region
        tuning_s += pretune(lats[0], bufs[0], gf.LANDAU, [args.mode], streams[0], sa_mode=args.sa_mode)
        _lib.call("gf_synchronize")
        t0 = time.perf_counter()  # the inputs were generated and uploaded before t0
        sa = gf.batch_run_sa(lats, bufs, gf.LANDAU, 1.45, 0.01, args.sa_steps, 3, args.reproject,
                             seeds=[1] * args.copies, counters=[(h - 10) * 100000 for h in hs], mode=args.sa_mode,
                             max_concurrent=args.copies, streams=streams, graph=graph)
        t1 = time.perf_counter()
        res = gf.batch_run_mixed(lats, bufs, gf.LANDAU, args.omega, args.switch, args.precision, args.maxiter,
                                 mode=args.mode, check32=args.check32, max_concurrent=args.copies,
                                 streams=streams)
        size_s = time.perf_counter() - t0
        total_s += size_s
        copies = []
        for c, ((_, ntemps), r) in enumerate(zip(sa, res)):
            total_su += (ntemps * (1 + 3) + r.sweeps) * V
            copies.append({"or_ms": r.ms_total, "or_sweeps": r.sweeps, "gff": r.gff, "theta": r.theta})
            print(f"bench: {L}^4 copy {c}: {r}", file=sys.stderr, flush=True)
        per_size[f"{L}^4"] = {"best_gff": max(x["gff"] for x in copies), "wall_s": size_s, "sa_s": t1 - t0,
                              "copies": copies}
    line = _base_line("Gribov copies: site-updates/s over SA + OR of 4 copies each of 16^4..32^4",
                      total_su / total_s, "site-updates/s", total_s * 1e3, 1, 0,
                      {"workload": "gribov-copies-16-20-24-32-sa-or", "sizes": sizes, "copies": args.copies,
                       "sa_graph": graph,
                       "sigma": 0.3, "seed": 1, "haar_seeds": [10, haar - 1], "sa": [1.45, 0.01, args.sa_steps, 3],
                       "precision": args.precision, "mode": MODE_NAMES[args.mode]})
    line.update({"total_s": total_s, "per_size": per_size, "tuning_s": tuning_s})
    return line


def gribov_one_launch_line(args, streams, graph=False, limit=None) -> dict:
    """The gauge copies of gribov_line (same configurations, seeds and RNG
    counters) with all sixteen annealed through one gf_batch_run_sa call with
    GF_SA_ONE_LAUNCH: each half-sweep of a group of lattices is one launch
    (groups of at most `limit` 64-site blocks per parity, the library's
    default 8192 when None; 0: all sixteen in one launch), then OR-fixed
    through one gf_batch_run_mixed call.  Every copy's result is that of
    gribov_line, bit for bit."""
    sizes = [16, 20, 24, 32]
    lats, bufs, hs_all, vols = [], [], [], []
    tuning_s, haar = 0.0, 10
    for L in sizes:
        dims = (L, L, L, L)
        lat = gf.Lattice(dims)
        hs = list(range(haar, haar + args.copies))
        haar += args.copies
        b = [gf.DeviceBuffer.from_numpy(gf.make_config(dims, 0.3, 1, h)) for h in hs]
        tuning_s += pretune(lat, b[0], gf.LANDAU, [args.mode], streams[0], sa_mode=args.sa_mode)
        lats += [lat] * args.copies
        bufs += b
        hs_all += hs
        vols += [L ** 4] * args.copies
    prev = gf.set_batch_launch_blocks(limit) if limit is not None else None
    try:
        _lib.call("gf_synchronize")
        t0 = time.perf_counter()
        sa = gf.batch_run_sa(lats, bufs, gf.LANDAU, 1.45, 0.01, args.sa_steps, 3, args.reproject,
                             seeds=[1] * len(lats), counters=[(h - 10) * 100000 for h in hs_all], mode=args.sa_mode,
                             max_concurrent=args.copies, streams=streams, graph=graph, one_launch=True)
        t1 = time.perf_counter()
        res = gf.batch_run_mixed(lats, bufs, gf.LANDAU, args.omega, args.switch, args.precision, args.maxiter,
                                 mode=args.mode, check32=args.check32, max_concurrent=args.copies, streams=streams)
        total_s = time.perf_counter() - t0
    finally:
        if prev is not None:
            gf.set_batch_launch_blocks(prev)
    per_size, total_su = {}, 0
    for k, ((_, ntemps), r) in enumerate(zip(sa, res)):
        total_su += (ntemps * (1 + 3) + r.sweeps) * vols[k]
        key = f"{sizes[k // args.copies]}^4"
        per_size.setdefault(key, {"best_gff": -1.0, "or_sweeps": []})
        per_size[key]["best_gff"] = max(per_size[key]["best_gff"], r.gff)
        per_size[key]["or_sweeps"].append(r.sweeps)
    print(f"bench: gauge copies, one-launch annealing (limit {limit}, graph {graph}): {total_s:.2f} s "
          f"(annealing {t1 - t0:.2f} s)", file=sys.stderr, flush=True)
    return {"total_s": total_s, "sa_s": t1 - t0, "or_s": total_s - (t1 - t0),
            "site_updates_per_s": total_su / total_s, "launch_block_limit": limit if limit is not None else 8192,
            "sa_graph": graph, "per_size": per_size, "tuning_s": tuning_s}


def small_copies_line(args, streams, L: int, n: int = 16) -> dict:
    """n gauge copies of an L^4 lattice (Haar seeds 100..), annealed as in
    gribov_line (args.sa_steps temperatures, 3 microcanonical steps each) once
    with a launch per copy (four copies concurrently, each on its own stream)
    and once with each half-sweep of all n copies in one launch
    (GF_SA_ONE_LAUNCH): lattices this small are bound by launch overhead, not
    by bytes.  Both runs must leave the same links."""
    dims = (L, L, L, L)
    lats = [gf.Lattice(dims)] * n
    cfgs = [gf.make_config(dims, 0.3, 1, 100 + c) for c in range(n)]
    kw = dict(seeds=[1] * n, counters=[c * 100000 for c in range(n)], mode=args.sa_mode, max_concurrent=4,
              streams=streams[:4])
    pretune(lats[0], gf.DeviceBuffer.from_numpy(cfgs[0]), gf.LANDAU, [], streams[0], sa_mode=args.sa_mode)
    out = {}
    for key, one in (("per_copy_launches_s", False), ("one_launch_s", True)):
        bufs = [gf.DeviceBuffer.from_numpy(c) for c in cfgs]
        _lib.call("gf_synchronize")
        t0 = time.perf_counter()
        gf.batch_run_sa(lats, bufs, gf.LANDAU, 1.45, 0.01, args.sa_steps, 3, args.reproject, one_launch=one, **kw)
        out[key] = time.perf_counter() - t0
        out[key + "_links"] = [b.download(np.float64, c.size) for b, c in zip(bufs, cfgs)]
    same = all(np.array_equal(a, b) for a, b in zip(out.pop("per_copy_launches_s_links"), out.pop("one_launch_s_links")))
    print(f"bench: {n} copies of {L}^4, annealing: {out['per_copy_launches_s']:.3f} s with a launch per copy, "
          f"{out['one_launch_s']:.3f} s in one launch", file=sys.stderr, flush=True)
    return {**out, "copies": n, "dims": list(dims), "identical_links": same,
            "speedup_one_launch": out["per_copy_launches_s"] / out["one_launch_s"]}


def same_copies(a: dict, b: dict) -> bool:
    """Whether two gauge-copies runs ended on the same best functional for
    every size (both runs take the same copies through the same updates)."""
    return all(a["per_size"][k]["best_gff"] == b["per_size"][k]["best_gff"] for k in a["per_size"])


def landau48_line(args) -> dict:
    """Landau gauge on a 48^4 lattice (sigma --sigma, seed 1, Haar seed 2):
    the working set (3.06 GB of FP64 links; 0.33 GB of 24-bit links and g in
    the FP32 phase) is far beyond the 256 MB Infinity Cache, so the kernels
    stream from HBM itself.  Mixed against FP64 to theta < --precision, and
    the dominant kernels' launch times with their fraction of 8 TB/s."""
    dims = (48, 48, 48, 48)
    V = int(np.prod(dims))
    U0 = gf.make_config(dims, args.sigma, 1, 2)
    lat64 = gf.Lattice(dims)
    lat32 = gf.Lattice(dims, gf.F32, gf.R12)
    src = gf.DeviceBuffer.from_numpy(U0)
    work = gf.DeviceBuffer(lat64.nbytes)
    stream = _stream()

    def reset():
        _lib.call("gf_memcpy_async", C.c_void_p(work.ptr), C.c_void_p(src.ptr), C.c_size_t(lat64.nbytes), stream)

    # a two-sweep run first: the library picks its wavefront slab for this
    # lattice on its first run (a few timed sweeps), outside the timed runs
    reset()
    gf.run_mixed(lat64, work, gf.LANDAU, args.omega, args.switch, args.precision, 2, mode=args.mode, stream=stream)
    reset()
    n_pipe = pipelined_launches()
    m = gf.run_mixed(lat64, work, gf.LANDAU, args.omega, args.switch, args.precision, args.maxiter, mode=args.mode,
                     refresh=args.refresh, check32=args.check32, reproject=args.reproject, stream=stream)
    n_pipe = pipelined_launches() - n_pipe
    print(f"bench: landau 48^4 mixed {m} ({n_pipe} pipelined sweeps)", file=sys.stderr, flush=True)
    slab, tpass = C.c_int(), C.c_int()
    _lib.call("gf_wavefront_choice", C.byref(lat64.struct), gf.LANDAU, args.mode, C.byref(slab))
    _lib.call("gf_block_order_choice", C.byref(lat64.struct), gf.LANDAU, args.mode, C.byref(tpass))
    # the same run with every sweep as two kernels (no wavefront)
    prev = C.c_int()
    _lib.call("gf_set_wavefront", 0, C.byref(prev))
    reset()
    m2 = gf.run_mixed(lat64, work, gf.LANDAU, args.omega, args.switch, args.precision, args.maxiter, mode=args.mode,
                      refresh=args.refresh, check32=args.check32, reproject=args.reproject, stream=stream)
    _lib.call("gf_set_wavefront", prev.value, None)
    print(f"bench: landau 48^4 mixed, two-kernel sweeps {m2}", file=sys.stderr, flush=True)
    reset()
    f = gf.run_fast(lat64, work, gf.LANDAU, args.omega, args.precision, args.maxiter, args.reproject, stream=stream)
    print(f"bench: landau 48^4 fp64 {f}", file=sys.stderr, flush=True)
    reset()
    fg = gf.run_mixed(lat64, work, gf.LANDAU, args.omega, args.switch, args.precision, args.maxiter,
                      mode=gf.MIXED_FP64, stream=stream)
    print(f"bench: landau 48^4 fp64 gauge sweep {fg}", file=sys.stderr, flush=True)
    U32 = gf.DeviceBuffer(lat32.nbytes)
    _lib.call("gf_convert", C.byref(lat64.struct), C.c_void_p(src.ptr), C.byref(lat32.struct), C.c_void_p(U32.ptr),
              0, stream)
    kms = {"gauge_sweep_q24": kernel_time_ms(lat32, U32, stream, args.kernel_sweeps, 4),
           "gauge_sweep_q24_wavefront": kernel_time_ms(lat32, U32, stream, args.kernel_sweeps, 6),
           "sweep_f64_tiled12": kernel_time_ms(lat64, src, stream, args.kernel_sweeps, 0),
           "gauge_sweep_f64_tiled12": kernel_time_ms(lat64, src, stream, args.kernel_sweeps, 7),
           "dropin_quality_f64_pattern18": kernel_time_ms(lat64, src, stream, args.kernel_sweeps, 5)}
    # per half-sweep (V/2 site-updates); the wavefront's figure is half a
    # whole sweep's time, so its fraction counts the algorithmic bytes of the
    # half-sweeps it replaces
    kbytes = {"gauge_sweep_q24": (V // 2) * 432, "gauge_sweep_q24_wavefront": (V // 2) * 432,
              "sweep_f64_tiled12": (V // 2) * 16 * 12 * 8, "gauge_sweep_f64_tiled12": (V // 2) * 1056,
              "dropin_quality_f64_pattern18": V * 4 * 12 * 8}
    both = m.converged == 1 and f.converged == 1
    return {
        "config": {"workload": "landau-or-48x48x48x48-mixed", "dims": list(dims), "sigma": args.sigma,
                   "precision": args.precision, "switch_precision": args.switch, "mode": MODE_NAMES[args.mode]},
        "time_to_theta_mixed_s": m.ms_total / 1e3, "sweeps_mixed": m.sweeps, "sweeps_fp32": m.sweeps_fp32,
        "converged_mixed": m.converged, "gff_mixed": m.gff, "theta_mixed": m.theta,
        "wavefront_slab": slab.value, "block_order_tpass": tpass.value, "sweep_path": sweep_path(slab.value),
        "pipelined_sweeps_in_mixed_run": n_pipe,
        "time_to_theta_mixed_two_kernel_sweeps_s": m2.ms_total / 1e3,
        "mixed_identical_with_and_without_wavefront": (m.sweeps, m.gff, m.theta) == (m2.sweeps, m2.gff, m2.theta),
        "time_to_theta_fp64_s": f.ms_total / 1e3, "sweeps_fp64": f.sweeps, "converged_fp64": f.converged,
        "gff_fp64": f.gff, "gff_rel_diff_mixed_vs_fp64": abs(m.gff - f.gff) / abs(f.gff),
        "speedup_mixed_vs_fp64": f.ms_total / m.ms_total if both else None,
        "fp64_gauge_sweep": {"time_to_theta_s": fg.ms_total / 1e3, "sweeps": fg.sweeps, "converged": fg.converged,
                             "gff_rel_diff_vs_fp64": abs(fg.gff - f.gff) / abs(f.gff),
                             "speedup_vs_fp64_link_sweeps": f.ms_total / fg.ms_total if fg.converged else None},
        "site_updates_per_s_mixed": m.sweeps * V / (m.ms_total / 1e3),
        "kernel_ms": kms,
        "kernel_hbm_frac": {k: kbytes[k] / (kms[k] / 1e3) / 1e9 / HBM_PEAK_GBS for k in kms},
    }


def extra_workloads(args, streams) -> dict:
    """The other workloads, one run each, for the default line: Coulomb
    32^3 x 64, the gauge copies (full schedule), MAG to theta < 1e-12 at 24^4
    (the 48^4 MAG run takes minutes: --workload mag) with MAG's kernels timed
    at 48^4, and Landau at 48^4 (HBM-resident)."""
    one = argparse.Namespace(**{**vars(args), "steps": 1, "warmup": 0})
    out = {}
    for name, fn in (("coulomb_32x32x32x64", lambda: coulomb_line(one)),
                     ("gauge_copies", lambda: gribov_line(one, streams)),
                     ("gauge_copies_sa_graph", lambda: gribov_line(one, streams, graph=True)),
                     ("gauge_copies_one_launch", lambda: gribov_one_launch_line(one, streams)),
                     ("mag_24x24x24x24", lambda: {**mag_line(one, 24, 200000), **mag_kernels(one, 48)}),
                     ("landau_48x48x48x48", lambda: landau48_line(one))):
        t0 = time.perf_counter()
        print(f"bench: workload {name}", file=sys.stderr, flush=True)
        out[name] = fn()
        out[name]["wall_s"] = time.perf_counter() - t0
        if name == "gauge_copies_one_launch":
            out[name]["same_best_gff_as_per_copy_launches"] = same_copies(out[name], out["gauge_copies"])
        _lib.call("gf_synchronize")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the library runs one lattice "
              "per GPU and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.workload != "landau":
        {"coulomb": workload_coulomb, "mag": workload_mag, "gribov": workload_gribov,
         "small-copies": workload_small_copies}[args.workload](args)
        return
    # the gauge copies' streams, created before the first device allocation
    # (streams created after it have run the concurrent copies up to 35% slower)
    copy_streams = None if args.no_extras else [_stream() for _ in range(args.copies)]
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

    def mixed(mode, switch):
        reset()
        return gf.run_mixed(lat64, work, gf.LANDAU, args.omega, switch, args.precision, args.maxiter, mode=mode,
                            refresh=args.refresh, check32=args.check32, reproject=args.reproject, stream=stream)

    def one_mixed():
        return mixed(args.mode, args.switch)

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

    # plain FP64 reference run (same driver, all sweeps FP64, on its tiled copy of the links)
    fp64 = None
    if not args.no_fp64:
        reset()
        fp64 = gf.run_fast(lat64, work, gf.LANDAU, args.omega, args.precision, args.maxiter, args.reproject,
                           stream=stream)
        print(f"bench: fp64 {fp64}", file=sys.stderr, flush=True)
    # all-FP64 on the FP64 gauge-transformation sweep (gf_run_mixed without
    # its FP32 phase): fewer bytes per site-update than the link sweep
    fp64g = None
    if not args.no_fp64:
        fp64g = mixed(gf.MIXED_FP64, args.switch)
        print(f"bench: fp64 gauge sweep {fp64g}", file=sys.stderr, flush=True)

    # the drop-in path end to end: the reference app's OR loop (gf_run_or:
    # reference-order kernels, the quality kernel and its reduction with a
    # host read after every sweep, projectSU3 every `reproject` sweeps)
    dropin = None
    if not args.no_fp64:
        reset()
        dropin = gf.run_or(lat64, work, gf.LANDAU, args.omega, args.precision, args.maxiter, 1, args.reproject,
                           stream=stream)
        print(f"bench: drop-in run_or {dropin}", file=sys.stderr, flush=True)

    # the same run with the FP32 -> FP64 switch at theta < 1e-6, and the other
    # single-precision modes
    ns = None
    modes = {}
    tuning_s = 0.0
    if not args.no_fp64:
        # the other modes' first runs would time their candidate sweep paths
        # inside their own timing (the warm-up did this for args.mode)
        tuning_s = pretune(lat64, src, gf.LANDAU, [m for m in (1, 2, 3) if m != args.mode], stream)
        ns = mixed(args.mode, args.early_switch)
        print(f"bench: mixed, switch {args.early_switch:g}: {ns}", file=sys.stderr, flush=True)
        for m in (0, 1, 2, 3):
            rm = r if m == args.mode else mixed(m, args.switch)
            modes[MODE_NAMES[m]] = {"time_to_theta_s": rm.ms_total / 1e3, "sweeps": rm.sweeps,
                                    "sweeps_fp32": rm.sweeps_fp32, "gff": rm.gff,
                                    "speedup_vs_fp64": None if fp64 is None else fp64.ms_total / rm.ms_total}
            print(f"bench: mode {MODE_NAMES[m]}: {rm}", file=sys.stderr, flush=True)

    # the tuner's choices for this lattice (gauge-sweep modes)
    tuning = None
    if args.mode in (1, 2, 3):
        slab, tpass = C.c_int(), C.c_int()
        _lib.call("gf_wavefront_choice", C.byref(lat64.struct), gf.LANDAU, args.mode, C.byref(slab))
        _lib.call("gf_block_order_choice", C.byref(lat64.struct), gf.LANDAU, args.mode, C.byref(tpass))
        tuning = {"wavefront_slab": slab.value, "block_order_tpass": tpass.value, "sweep_path": sweep_path(slab.value)}

    # kernel timing for the roofline (on the launch stream)
    U32 = gf.DeviceBuffer(lat32.nbytes)

    def fresh32():
        _lib.call("gf_convert", C.byref(lat64.struct), C.c_void_p(src.ptr), C.byref(lat32.struct),
                  C.c_void_p(U32.ptr), 0, stream)

    kms = {}
    for var, name in ((1, "sweep_f32_r12_gauge"), (2, "gauge_sweep_f32"), (3, "gauge_sweep_f64arith"),
                      (4, "gauge_sweep_q24")):
        fresh32()
        kms[name] = kernel_time_ms(lat32, U32, stream, args.kernel_sweeps, var)
    reset()
    # the FP64 kernels on the drivers' tiled copy of the links: the link sweep
    # (gf_run_fast) and the gauge-transformation sweep (the FP64 phase of the
    # gauge modes)
    kms["sweep_f64_tiled12"] = kernel_time_ms(lat64, work, stream, args.kernel_sweeps, 0)
    kms["gauge_sweep_f64_tiled12"] = kernel_time_ms(lat64, work, stream, args.kernel_sweeps, 7)
    reset()
    kms["dropin_sweep_f64_pattern18"] = dropin_kernel_ms(lat64, work, stream, args.kernel_sweeps)
    # the drop-in quality kernel (GaugeFixingStats::generateGaugeQuality's per-site pass)
    kms["dropin_quality_f64_pattern18"] = kernel_time_ms(lat64, work, stream, args.kernel_sweeps, 5)
    # algorithmic bytes per parity launch (V/2 site-updates):
    #   link sweeps: 16 R sizeof(real), R = 12 (rows 0,1 of every link read and written once);
    #   FP32 links + g: also the site's 12-float g read and written (+96 B);
    #   gauge-transformation sweep: U0 read once (8 x 48 B), g of the other parity read once
    #   (48 B), the site's g read and written (96 B): 528 B; with U0 in 24-bit fixed
    #   point (8 x 36 B): 432 B; all in FP64 (links 8 x 96 B, g 96 + 192 B): 1056 B
    kbytes = {
        "sweep_f32_r12_gauge": (V // 2) * (16 * 12 * 4 + 2 * 12 * 4),
        "gauge_sweep_f32": (V // 2) * 528,
        "gauge_sweep_f64arith": (V // 2) * 528,
        "gauge_sweep_q24": (V // 2) * 432,
        "sweep_f64_tiled12": (V // 2) * 16 * 12 * 8,
        "gauge_sweep_f64_tiled12": (V // 2) * 1056,
        "dropin_sweep_f64_pattern18": (V // 2) * 16 * 12 * 8,
        # quality: every link's rows 0,1 read once per launch (V sites x 4 links x 12 reals)
        "dropin_quality_f64_pattern18": V * 4 * 12 * 8,
    }
    t32 = r.ms_fp32 / max(r.sweeps_fp32, 1)
    t64 = (r.ms_total - r.ms_fp32) / max(r.sweeps - r.sweeps_fp32, 1)
    k32name = {0: "sweep_f32_r12_gauge", 1: "gauge_sweep_f32", 2: "gauge_sweep_f64arith",
               3: "gauge_sweep_q24"}[args.mode]
    k64name = "sweep_f64_tiled12" if args.mode == 0 else "gauge_sweep_f64_tiled12"
    dom32 = r.sweeps_fp32 * kms[k32name] >= (r.sweeps - r.sweeps_fp32) * kms[k64name]
    kname = k32name if dom32 else k64name
    achieved = kbytes[kname] / (kms[kname] / 1e3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(kname, {}).get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None

    cpu = cpu1 = None
    if not args.no_cpu:
        cpu = cpu_baseline(dims, U0, args.cpu_seconds)
        cpu1 = cpu_baseline(dims, U0, args.cpu_seconds / 4, threads=1)

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
            "mode": MODE_NAMES[args.mode],
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
        "speedup_mixed_vs_fastest_fp64": None if fp64 is None or fp64g is None
        else min(fp64.ms_total, fp64g.ms_total) / r.ms_total,
        "fp64_gauge_sweep": None if fp64g is None else {
            "time_to_theta_s": fp64g.ms_total / 1e3, "sweeps": fp64g.sweeps, "gff": fp64g.gff,
            "theta": fp64g.theta, "gff_rel_diff_vs_fp64": abs(fp64g.gff - fp64.gff) / abs(fp64.gff),
            "speedup_vs_fp64_link_sweeps": fp64.ms_total / fp64g.ms_total},
        "mixed_switch_1e-6": None if ns is None else {
            "time_to_theta_s": ns.ms_total / 1e3, "sweeps": ns.sweeps, "sweeps_fp32": ns.sweeps_fp32,
            "gff": ns.gff, "speedup_vs_fp64": fp64.ms_total / ns.ms_total},
        "mixed_modes": modes,
        "mixed_modes_tuning_s": tuning_s,
        "dropin_run_or": None if dropin is None else {
            "time_to_theta_s": dropin.ms_total / 1e3, "sweeps": dropin.sweeps, "gff": dropin.gff,
            "theta": dropin.theta, "site_updates_per_s": dropin.sweeps * V / (dropin.ms_total / 1e3)},
        "site_updates_per_s_fp64": None if fp64 is None else fp64.sweeps * V / (fp64.ms_total / 1e3),
        "sweep_ms_in_run": {"fp32": t32, "fp64": t64},
        "tuning": tuning,
        "kernel_ms": kms,
        "kernel_hbm_frac": {k: kbytes[k] / (kms[k] / 1e3) / 1e9 / HBM_PEAK_GBS for k in kms},
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": kbytes[kname],
        },
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "wall_s": wall,
    }
    _lib.call("gf_stream_destroy", stream)
    if not args.no_extras:
        del src, work, U32
        line["workloads"] = extra_workloads(args, copy_streams)
        # the same kernel's fraction where the lattice streams from HBM (48^4,
        # 0.33 GB of 24-bit links and g beyond the 256 MB Infinity Cache): a
        # whole sweep without a check as the mixed run makes it, per half-sweep
        w48 = line["workloads"].get("landau_48x48x48x48", {})
        k48 = w48.get("kernel_ms", {}).get("gauge_sweep_q24_wavefront")
        if k48:
            a48 = (48 ** 4 // 2) * 432 / (k48 / 1e3) / 1e9
            line["roofline_hbm_resident"] = {
                "bound": "hbm", "kernel": "gauge_sweep_q24, whole sweeps (" + w48.get("sweep_path", "?") + ")",
                "dims": [48, 48, 48, 48], "achieved": a48, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": a48 / HBM_PEAK_GBS, "algorithmic_bytes_per_half_sweep": (48 ** 4 // 2) * 432,
                "ms_per_half_sweep": k48}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
