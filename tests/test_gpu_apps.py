"""The gauge-fixing application on the drop-in C++ headers (include/culgt),
run as a separate process on the GPU and compared RESULT line by RESULT line
with the same schedule (tests/app_schedule.py) run

* through the library's Python mirror on the GPU (GpuBackend): the same
  kernels in the same order with the same RNG counters, so every result must
  be identical -- this pins the application's control flow exactly, heatbath
  schedules included;
* on the CPU oracle: exact sweep counts and RNG counters, and the functional
  to the reduction order (1e-13) for OR-only schedules, to 1e-9 where a
  heatbath or random transformation precedes OR at low temperature.  At
  higher temperatures the heatbath amplifies last-bit differences of log /
  cos / sin (device vs host library) into different trajectories; there only
  the counters and the schedule's length are compared with the oracle.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from app_schedule import GpuBackend, Options, parse_results, run_4d, run_coulomb

BIN = Path(__file__).resolve().parent.parent / "su3_ma_gauge_gpu_amd" / "apps" / "bin"
DIMS = (8, 8, 8, 8)  # the extents the x8t8 binaries are compiled for
CFG = dict(sigma=0.3, cfgseed=1, haar=2)

pytestmark = pytest.mark.gpu


def run_app(name: str, gauge: str, o: Options, timeout: int = 100) -> tuple[list[dict], str]:
    exe = BIN / name
    assert exe.exists(), f"{exe} not built (python -m su3_ma_gauge_gpu_amd.build)"
    args = [str(exe), "--gauge", gauge, *o.argv(), "--sigma", str(CFG["sigma"]), "--cfgseed", str(CFG["cfgseed"]),
            "--haar", str(CFG["haar"])]
    print("app:", " ".join(args[1:]), flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    return parse_results(res.stdout), res.stdout


def config(gf, dtype=np.float64):
    return gf.make_config(DIMS, CFG["sigma"], CFG["cfgseed"], CFG["haar"]).astype(dtype)


def same(got, want):
    """Identical results: the application and the library ran the same kernels."""
    assert len(got) == len(want)
    for g, w in zip(got, want):
        for k in ("gff", "theta", "or_sweeps", "sr_sweeps", "counter", "slice", "copy"):
            if k in g and k in w:
                assert g[k] == w[k], (k, g, w)


def compare(got, want, rel_gff, exact_sweeps=True):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g["gff"] == pytest.approx(w["gff"], rel=rel_gff), (g, w)
        for k in ("or_sweeps", "sr_sweeps", "counter"):
            if k in g and k in w and (exact_sweeps or k == "counter"):
                assert g[k] == w[k], (k, g, w)


def test_landau_or_only_f64(gf, gpu, oracle):
    """The reference's OR loop alone (no random transformation, no SA):
    bit-exact sweeps, so the stopping sweep is the oracle's."""
    o = Options(randomtrafo=0, sasteps=0, precision=1e-12, ormaxiter=5000, checkprecision=10)
    got, _ = run_app("gaugefix_x8t8_f64", "landau", o)
    want = run_4d(oracle, "landau", DIMS, config(gf), o)
    compare(got, want, 1e-13)
    assert got[0]["theta"] < 1e-12
    assert got[0]["theta"] == pytest.approx(want[0]["theta"], rel=1e-9)


def test_landau_or_only_f32(gf, gpu, oracle32):
    """Single precision (the reference's default build, Real = float)."""
    o = Options(randomtrafo=0, sasteps=0, precision=1e-6, ormaxiter=3000, checkprecision=10)
    got, _ = run_app("gaugefix_x8t8_f32", "landau", o)
    want = run_4d(oracle32, "landau", DIMS, config(gf, np.float32), o)
    compare(got, want, 1e-6)
    assert got[0]["theta"] < 1e-6


def test_landau_or_only_f32_dp_updates(gf, gpu, oracle32dp):
    """The single-precision build with DPUPDATES=true (-DUSE_DP_ORUPDATE
    -DUSE_DP_MICROUPDATE -DUSE_DP_SAUPDATE, apps/Makefile:46-48): the drop-in
    headers select the double-precision update rules, and the OR loop stops
    on the oracle's sweep with its functional."""
    o = Options(randomtrafo=0, sasteps=0, precision=1e-6, ormaxiter=3000, checkprecision=10)
    got, _ = run_app("gaugefix_x8t8_f32dp", "landau", o)
    want = run_4d(oracle32dp, "landau", DIMS, config(gf, np.float32), o)
    compare(got, want, 1e-6)
    assert got[0]["theta"] < 1e-6


def test_landau_sa_copies_f64(gf, gpu, oracle):
    """Random transformation, 30 SA temperatures with 3 microcanonical steps
    each, then OR to 1e-12, for two gauge copies."""
    o = Options(gaugecopies=2, sasteps=30, precision=1e-12, ormaxiter=5000, checkprecision=20)
    got, out = run_app("gaugefix_x8t8_f64", "landau", o)
    same(got, run_4d(GpuBackend(gf), "landau", DIMS, config(gf), o))
    want = run_4d(oracle, "landau", DIMS, config(gf), o)
    compare(got, want, 1e-9)
    best = [float(line.split("=")[1]) for line in out.splitlines() if line.startswith("BEST")]
    assert best == [pytest.approx(max(w["gff"] for w in want), rel=1e-9)]


def test_landau_hot_start_f64(gf, gpu, oracle):
    """--hotgaugefield: setHot once, the copies continue from each other.
    OR only: far from a maximum, heatbath steps would amplify the last-bit
    differences of setHot's transcendentals beyond any fixed tolerance."""
    o = Options(gaugecopies=2, hotgaugefield=1, randomtrafo=0, sasteps=0, precision=1e-10, ormaxiter=300,
                checkprecision=20)
    got, _ = run_app("gaugefix_x8t8_f64", "landau", o)
    zero = np.zeros(DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3] * 72)
    same(got, run_4d(GpuBackend(gf), "landau", DIMS, zero, o))
    compare(got, run_4d(oracle, "landau", DIMS, zero, o), 1e-9)


def test_u1xu1_fine_steps_f64(gf, gpu, oracle):
    """The U1xU1 annealing schedule with its fine steps for 2 <= T <= 2.7:
    identical to the library-driven schedule; the number of temperatures
    visited (two RNG counters each) and the OR sweep count are the oracle's."""
    o = Options(samax=2.8, samin=1.9, sasteps=10, ormaxiter=150, checkprecision=50)
    got, _ = run_app("gaugefix_x8t8_f64", "u1xu1", o)
    same(got, run_4d(GpuBackend(gf), "u1xu1", DIMS, config(gf), o))
    want = run_4d(oracle, "u1xu1", DIMS, config(gf), o)
    # two counters per temperature: > 100 temperatures are visited only with the fine steps
    assert got[0]["counter"] == want[0]["counter"] > 200
    assert got[0]["or_sweeps"] == want[0]["or_sweeps"] == 150


def test_u1xu1_cold_anneal_f64(gf, gpu, oracle):
    """U1xU1 with two heatbath temperatures near zero (the U1xU1 app always
    anneals at least once), no random transformation, then OR: the heatbath
    is then all but deterministic and the functional is the oracle's to 1e-9."""
    o = Options(randomtrafo=0, sasteps=1, samax=0.01, samin=0.005, microupdates=0, ormaxiter=150,
                checkprecision=50)
    got, _ = run_app("gaugefix_x8t8_f64", "u1xu1", o)
    same(got, run_4d(GpuBackend(gf), "u1xu1", DIMS, config(gf), o))
    compare(got, run_4d(oracle, "u1xu1", DIMS, config(gf), o), 1e-9)


def test_mag_schedule_f64(gf, gpu, oracle):
    """MAG: five heatbath steps per temperature, SR until theta < 1e-6, then
    OR: identical to the library-driven schedule; counters and sweep counts
    are the oracle's."""
    o = Options(sasteps=6, srmaxiter=400, ormaxiter=600, precision=1e-8, checkprecision=20)
    got, _ = run_app("gaugefix_x8t8_f64", "mag", o)
    same(got, run_4d(GpuBackend(gf), "mag", DIMS, config(gf), o))
    want = run_4d(oracle, "mag", DIMS, config(gf), o)
    compare(got, want, 1e-3)
    assert got[0]["sr_sweeps"] > 1


def test_mag_sr_or_f64(gf, gpu, oracle):
    """MAG without annealing or random transformation (SR, then OR): the SR
    updates draw uniform numbers only (no log / cos / sin), so the run is the
    oracle's to 1e-9."""
    o = Options(randomtrafo=0, doSA=0, srmaxiter=200, ormaxiter=300, precision=1e-8, checkprecision=20)
    got, _ = run_app("gaugefix_x8t8_f64", "mag", o)
    same(got, run_4d(GpuBackend(gf), "mag", DIMS, config(gf), o))
    compare(got, run_4d(oracle, "mag", DIMS, config(gf), o), 1e-9)


def test_coulomb_slices_f64(gf, gpu, oracle):
    """Coulomb: every timeslice in turn, with slice t-1's time links."""
    o = Options(sasteps=4, precision=1e-12, ormaxiter=3000, checkprecision=20)
    got, _ = run_app("gaugefix_x8t8_f64", "coulomb", o)
    same(got, run_coulomb(GpuBackend(gf), DIMS, config(gf), o))
    want = run_coulomb(oracle, DIMS, config(gf), o)
    assert [g["slice"] for g in got] == list(range(DIMS[0]))
    compare(got, want, 1e-9)
