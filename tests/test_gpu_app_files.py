"""The application's file handling (the reference apps' FileIterator loop with
LinkFile / ILDG / QCDSTAG load and save) on the GPU: a configuration written
to a file in each format and gauge-fixed by the application gives the same
results as the application on the same configuration from the generator, the
saved file holds the best copy (its functional, by the oracle, is the one the
application printed), headers and ILDG records are carried over, and MAG's
per-copy functional file and save_each follow the reference's rules."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import lime
from app_schedule import Options, parse_results
from oracle_lib import LANDAU, MAG

BIN = Path(__file__).resolve().parent.parent / "su3_ma_gauge_gpu_amd" / "apps" / "bin"
DIMS = (8, 8, 8, 8)
V = int(np.prod(DIMS))
CFG = dict(sigma=0.3, cfgseed=1, haar=2)
EXT = {"PLAIN": ".plain", "HEADERONLY": ".hdr", "VOGT": ".vogt", "ILDG": ".ildg", "QCDSTAG": ".q"}

pytestmark = pytest.mark.gpu


def run(gauge: str, o: Options, extra: list[str], timeout: int = 100) -> str:
    exe = BIN / "gaugefix_x8t8_f64"
    assert exe.exists(), f"{exe} not built (python -m su3_ma_gauge_gpu_amd.build)"
    args = [str(exe), "--gauge", gauge, *o.argv(), *extra]
    print("app:", " ".join(args[1:]), flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    return res.stdout


def synthetic(gauge: str, o: Options) -> str:
    return run(gauge, o, ["--sigma", str(CFG["sigma"]), "--cfgseed", str(CFG["cfgseed"]), "--haar", str(CFG["haar"])])


def write_input(gf, ftype: str, path: Path, U: np.ndarray, pattern: int = 0) -> None:
    kind = getattr(gf, f"FILE_{ftype}")
    f = gf.LinkFile(kind, pattern=pattern)
    if ftype == "ILDG":
        tmpl = path.with_suffix(".template")
        tmpl.write_bytes(lime.message([("xlf-info", True, False, b"ensemble A\0"),
                                       ("ildg-binary-data", False, True, bytes(8 * 72 * V))]))
        f.read_header(DIMS, tmpl)
    f.save(DIMS, path, U)
    if ftype == "HEADERONLY":
        path.write_bytes(b"# an mdp-like header\n" + path.read_bytes())


def best(out: str) -> list[float]:
    return [float(m) for m in re.findall(r"^BEST gff=(\S+)", out, flags=re.M)]


@pytest.mark.parametrize("ftype", ["VOGT", "ILDG", "QCDSTAG", "HEADERONLY", "PLAIN"])
def test_landau_files(gf, gpu, oracle, tmp_path, ftype):
    U = gf.make_config(DIMS, CFG["sigma"], CFG["cfgseed"], CFG["haar"])
    o = Options(gaugecopies=2, sasteps=20, ormaxiter=400, precision=1e-11, checkprecision=10)
    want = parse_results(synthetic("landau", o))
    ext = EXT[ftype]
    for i in (3, 5):  # fstartnumber 3, fstepnumber 2, nconf 2
        write_input(gf, ftype, tmp_path / f"cfg_{i:04d}{ext}", U)
    (tmp_path / "out").mkdir()
    out = run("landau", o, ["--ftype", ftype.lower(), "--fbasename", str(tmp_path / "cfg_"), "--fending", ext,
                            "--fnumberformat", "4", "--fstartnumber", "3", "--fstepnumber", "2", "--nconf", "2",
                            "--output_conf", str(tmp_path / "out") + "/", "--fappendix", "fixed_",
                            "--output_ending", ext])
    got = parse_results(out)
    assert len(got) == 4  # two files, two copies each
    # the first file starts from the same RNG counter as the generator's run
    for g, w in zip(got[:2], want):
        for k in ("gff", "theta", "or_sweeps", "counter"):
            assert g[k] == w[k], (k, g, w)
    for i, b in zip((3, 5), best(out)):
        p = tmp_path / "out" / f"fixed_{i:04d}{ext}"
        Uo = np.zeros(72 * V)
        f = gf.LinkFile(getattr(gf, f"FILE_{ftype}"))
        f.load(DIMS, p, Uo)
        g, th = oracle.quality(LANDAU, DIMS, Uo.copy())
        assert g == pytest.approx(b, rel=1e-12) and th < 1e-10
    saved = (tmp_path / "out" / f"fixed_0003{ext}").read_bytes()
    if ftype == "VOGT":
        assert saved[:14] == (tmp_path / f"cfg_0003{ext}").read_bytes()[:14]
    if ftype == "HEADERONLY":
        assert saved.startswith(b"# an mdp-like header\n")
    if ftype == "ILDG":
        # the input was written with 0 annealing steps; the app appends its own
        assert lime.parse(saved)[0][3] == b"ensemble A SA steps 0 SA steps 20"


def test_coulomb_qcdstag_timeslices(gf, gpu, tmp_path):
    U = gf.make_config(DIMS, CFG["sigma"], CFG["cfgseed"], CFG["haar"])
    o = Options(gaugecopies=1, sasteps=10, ormaxiter=300, precision=1e-11, checkprecision=10)
    want = parse_results(synthetic("coulomb", o))
    # the Coulomb app reads into the timeslice layout: the same file as the 4D one
    gf.LinkFile(gf.FILE_QCDSTAG).save(DIMS, tmp_path / "cfg_0.q", U)
    out = run("coulomb", o, ["--ftype", "QCDSTAG", "--fbasename", str(tmp_path / "cfg_"), "--fending", ".q",
                             "--output_conf", str(tmp_path) + "/", "--fappendix", "fixed_"])
    got = parse_results(out)
    assert len(got) == len(want) == DIMS[0]
    for g, w in zip(got, want):
        for k in ("slice", "gff", "theta", "or_sweeps", "counter"):
            assert g[k] == w[k], (k, g, w)
    Ut = np.zeros(72 * V)
    gf.LinkFile(gf.FILE_QCDSTAG, pattern=gf.PATTERN_GPU_TIMESLICE).load(DIMS, tmp_path / "fixed_0", Ut)
    U4 = np.zeros(72 * V)
    gf.LinkFile(gf.FILE_QCDSTAG).load(DIMS, tmp_path / "fixed_0", U4)
    assert not np.array_equal(U4, U)  # fixed, and the same file read in either layout
    from app_schedule import to_timeslices
    assert np.array_equal(to_timeslices(U4, DIMS).reshape(-1), Ut)


def test_mag_functional_file_and_save_each(gf, gpu, oracle, tmp_path):
    U = gf.make_config(DIMS, 0.35, 4, 5)
    gf.LinkFile(gf.FILE_VOGT).save(DIMS, tmp_path / "cfg_0.vogt", U)
    o = Options(gaugecopies=3, sasteps=8, samax=0.4, samin=0.35, srmaxiter=50, ormaxiter=100, precision=1e-9,
                checkprecision=10)
    out = run("mag", o, ["--ftype", "VOGT", "--fbasename", str(tmp_path / "cfg_"), "--output_conf",
                         str(tmp_path) + "/", "--output_SA_functional", str(tmp_path / "func"), "--save_each", "1"])
    got = parse_results(out)
    assert len(got) == 3
    lines = (tmp_path / "func_0").read_text().splitlines()
    assert lines[0] == "copy,functional" and len(lines) == 4
    # each better copy but the last is saved as <output>_<copy + 1>; no final save
    gffs = [r["gff"] for r in got]
    best_so_far, expect = 0.0, []
    for c, g in enumerate(gffs):
        if g > best_so_far:
            best_so_far = g
            if c < 2:
                expect.append(c + 1)
    for k in (1, 2, 3):
        assert (tmp_path / f"gaugefixed_0_{k}").exists() == (k in expect), (k, expect)
    assert not (tmp_path / "gaugefixed_0").exists()
    for k in expect:
        Uo = np.zeros(72 * V)
        gf.LinkFile(gf.FILE_VOGT).load(DIMS, tmp_path / f"gaugefixed_0_{k}", Uo)
        assert oracle.quality(MAG, DIMS, Uo.copy())[0] == pytest.approx(gffs[k - 1], rel=1e-12)


def test_config_file_and_overrides(gf, gpu, tmp_path):
    """A config file of "key = value" lines (comments after '#'), positional or
    --config-file, with the command line taking precedence
    (ProgramOptions.hxx:221-286): the same results as all options on the
    command line."""
    o = Options(gaugecopies=1, sasteps=10, ormaxiter=200, precision=1e-10, checkprecision=10)
    want = parse_results(synthetic("landau", o))
    cfg = tmp_path / "run.cfg"
    cfg.write_text("# gauge fixing\nsasteps = 10\normaxiter=200   # sweeps\nprecision = 1e-10\n"
                   "checkprecision = 10\nsigma = 0.3\ncfgseed = 1\nhaar = 2\nseed = 99\n")
    exe = BIN / "gaugefix_x8t8_f64"
    for how in (["--seed", "1", str(cfg)], ["--config-file", str(cfg), "--seed", "1"]):
        res = subprocess.run([str(exe), "--gauge", "landau", *how], capture_output=True, text=True, timeout=100)
        assert res.returncode == 0, res.stderr[-2000:]
        got = parse_results(res.stdout)
        assert [(g["gff"], g["or_sweeps"], g["counter"]) for g in got] == [(w["gff"], w["or_sweeps"], w["counter"])
                                                                           for w in want]
