"""The drop-in file headers (include/culgt: LinkFile, the ILDG and QCDSTAG
functions, the patterns and SiteCoord) in a C++ program built the way an app
is built, against the reference's own code (oracle/_ref/libref_io.so) and the
library's file handle.  CPU only: the file code never touches a GPU."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import lime

ROOT = Path(__file__).resolve().parent.parent
DIMS = (2, 4, 6, 8)
V = int(np.prod(DIMS))


@pytest.fixture(scope="module")
def exe(tmp_path_factory, gf):
    out = tmp_path_factory.mktemp("files") / "linkfile_check"
    lib = ROOT / "su3_ma_gauge_gpu_amd"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-D_T_=2", "-D_X_=4", "-D_Y_=6", "-D_Z_=8", "-DDOUBLEPRECISION", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "culgt" / "linkfile_check.cpp"), "-o", str(out), f"-L{lib}", "-lsu3gf",
                    f"-Wl,-rpath,{lib}"], check=True)
    return out


def test_pattern_index_by_unique(exe):
    out = subprocess.run([str(exe), "index"], capture_output=True, text=True, check=True).stdout
    got = np.array(out.split(), dtype=np.int64).reshape(-1, 2)
    assert got.shape == (72 * V, 2)
    # GpuPattern: a permutation onto the array; GpuPatternTimeslice likewise
    for col in (0, 1):
        assert np.array_equal(np.sort(got[:, col]), np.arange(72 * V))
    from ref_io import PATH, RefIO
    if PATH.exists():
        ref = RefIO()
        for u in range(0, 72 * V, 7):
            assert got[u, 0] == ref.index_by_unique(DIMS, u, 0)
            assert got[u, 1] == ref.index_by_unique(DIMS, u, 1)


def test_files_through_the_headers(exe, gf, tmp_path):
    U = np.random.default_rng(4).standard_normal(72 * V)
    raw = tmp_path / "U.raw"
    U.tofile(raw)
    recs = [("xlf-info", True, False, b"beta 6.0\0"), ("ildg-binary-data", False, True, bytes(8 * 72 * V))]
    (tmp_path / "template.lime").write_bytes(lime.message(recs))
    subprocess.run([str(exe), "write", str(raw), str(tmp_path)], check=True)
    # the same files through the library's handle directly
    for name, kind in (("plain", gf.FILE_PLAIN), ("vogt", gf.FILE_VOGT), ("qcdstag", gf.FILE_QCDSTAG)):
        p = tmp_path / f"lib_{name}"
        gf.LinkFile(kind).save(DIMS, p, U)
        assert (tmp_path / name).read_bytes() == p.read_bytes(), name
    p = tmp_path / "lib_qcdstag_ts"
    gf.LinkFile(gf.FILE_QCDSTAG, pattern=gf.PATTERN_GPU_TIMESLICE).save(DIMS, p, U)
    assert (tmp_path / "qcdstag_ts").read_bytes() == p.read_bytes()
    assert lime.parse((tmp_path / "ildg").read_bytes())[0][3] == b"beta 6.0 SA steps 5"
    from ref_io import PATH, RefIO
    if PATH.exists():
        ref = RefIO()
        ref.write_ildg(DIMS, tmp_path / "template.lime", tmp_path / "ref_ildg", U, 5)
        assert (tmp_path / "ildg").read_bytes() == (tmp_path / "ref_ildg").read_bytes()
        ref.write_qcdstag(DIMS, tmp_path / "ref_q", U)
        assert (tmp_path / "qcdstag").read_bytes() == (tmp_path / "ref_q").read_bytes()
    # read back, and a HEADERONLY file with a header of its own
    (tmp_path / "headeronly").write_bytes(b"HDR 7 bytes" + (tmp_path / "plain").read_bytes())
    out = subprocess.run([str(exe), "read", str(raw), str(tmp_path)], capture_output=True, text=True, check=True).stdout
    for fmt in ("plain", "vogt", "headeronly", "qcdstag", "ildg", "missing", "options"):
        assert f"ok {fmt}" in out, (fmt, out)
    assert (tmp_path / "headeronly_out").read_bytes() == (tmp_path / "headeronly").read_bytes()


def test_app_help_lists_the_reference_options():
    exe = ROOT / "su3_ma_gauge_gpu_amd" / "apps" / "bin" / "gaugefix_x8t8_f64"
    if not exe.exists():
        pytest.skip("apps not built")
    res = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1
    for opt in ("--ftype", "--fbasename", "--nconf", "--reinterpret", "--output_conf", "--save_each",
                "--devicenumber", "--config-file", "--checkprecision"):
        assert opt in res.stdout, opt
