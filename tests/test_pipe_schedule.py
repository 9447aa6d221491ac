"""The pipelined sweep's schedule (su3_ma_gauge_gpu_amd/csrc/su3gf_pipe.h),
checked on the CPU site by site by tests/pipe_schedule_check.cpp: each
queue's tickets give each even and odd 64-site block once, and every even
block holding a neighbour of an odd block's site is inside that odd block's
wait set and comes at an earlier ticket -- for no extra lag and for several.
The checker itself is shown to catch a lag one timeslice-set too short and a
neighbour reach rounded down."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "pipe_schedule_check.cpp"
HEADER = ROOT / "su3_ma_gauge_gpu_amd" / "csrc" / "su3gf_pipe.h"

SHAPES = [(8, 8, 8, 8), (16, 16, 16, 16), (8, 4, 8, 16), (16, 8, 12, 16), (24, 24, 24, 24), (32, 32, 32, 32),
          (40, 40, 40, 40), (48, 48, 48, 48), (16, 32, 32, 32), (8, 16, 16, 16), (64, 32, 32, 32), (24, 8, 8, 8)]


def _build(tmp_path, header_text=None):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src_dir = tmp_path / "tests"
    hdr_dir = tmp_path / "su3_ma_gauge_gpu_amd" / "csrc"
    src_dir.mkdir(parents=True)
    hdr_dir.mkdir(parents=True)
    (src_dir / SRC.name).write_text(SRC.read_text())
    (hdr_dir / HEADER.name).write_text(header_text if header_text is not None else HEADER.read_text())
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", str(src_dir / SRC.name), "-o", str(exe)], check=True)
    return exe


def _run(exe, shapes):
    args = [str(x) for s in shapes for x in s]
    return subprocess.run([str(exe), *args], capture_output=True, text=True)


def test_schedule_dependencies_come_first(tmp_path):
    res = _run(_build(tmp_path), SHAPES)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("ok ") == 5 * len(SHAPES)


@pytest.mark.parametrize("old,new,shape", [
    ("const int64_t L = (int64_t)(2 * p.R + 1) * p.T", "const int64_t L = (int64_t)(2 * p.R) * p.T", (32, 32, 32, 32)),
    ("p.R = (int)((63 + (int64_t)N[2] * N[3] / 2) / 64);", "p.R = (int)((int64_t)N[2] * N[3] / 2 / 64);",
     (16, 8, 12, 16)),
])
def test_checker_catches_planted_mistakes(tmp_path, old, new, shape):
    text = HEADER.read_text()
    assert old in text
    res = _run(_build(tmp_path, text.replace(old, new)), [shape])
    assert res.returncode == 1 and "FAIL" in res.stdout, res.stdout
