"""Configuration files (include/su3gf_io.h) on the CPU: the library's readers
and writers against the reference's own code.

Pins:
1. tests/golden/io_vectors.json: SHA-256 digests of files written by the
   reference's writeILDG / writeQCDSTAG (and their timeslice forms), its
   c-lime writer and a VOGT file from its FileVogt header and LinkFile's
   standard order (tests/golden/make_io_golden.py).  The library's writer must
   produce the same bytes; the LIME restatement tests/lime.py must produce the
   same template.
2. Where `make -C oracle ref` has built oracle/_ref/libref_io.so: the same
   comparisons live at a lattice with four different extents, the
   reference's readers on the library's files and the library's readers on the
   reference's files, c-lime parsing the library's LIME output, and the
   standard-order index map against the reference's pattern classes.
3. Round trips for every format, reinterpretation, precision and memory
   pattern, and the error paths.
"""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

import lime

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "io_vectors.json").read_text())
FORMATS = ["FILE_PLAIN", "FILE_HEADERONLY", "FILE_VOGT", "FILE_ILDG", "FILE_QCDSTAG"]


def sha(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def gold_links():
    V = int(np.prod(GOLD["dims"]))
    return np.random.default_rng(GOLD["seed"]).standard_normal(72 * V)


def gold_template():
    recs = []
    for t, mb, me, d in GOLD["template"]:
        recs.append((t, mb, me, bytes(d) if isinstance(d, int) else bytes.fromhex(d)))
    return recs


@pytest.fixture(scope="module")
def refio():
    from ref_io import PATH, RefIO
    if not PATH.exists():
        pytest.skip("oracle/_ref/libref_io.so not built (needs the reference sources)")
    return RefIO()


# ------------------------------------------------------------ fixtures
def test_lime_restatement_matches_c_lime_template(tmp_path):
    p = tmp_path / "template.lime"
    p.write_bytes(lime.message(gold_template()))
    assert sha(p) == GOLD["sha256"]["template"]


@pytest.mark.parametrize("timeslice", [0, 1])
def test_ildg_writer_matches_reference_bytes(gf, tmp_path, timeslice):
    dims, U = tuple(GOLD["dims"]), gold_links()
    tmpl = tmp_path / "template.lime"
    tmpl.write_bytes(lime.message(gold_template()))
    f = gf.LinkFile(gf.FILE_ILDG, pattern=timeslice)
    f.read_header(dims, tmpl)
    out = tmp_path / "out.ildg"
    f.save(dims, out, U, sa_steps=GOLD["sa_steps"])
    assert sha(out) == GOLD["sha256"]["ildg_timeslice" if timeslice else "ildg"]
    back = np.zeros_like(U)
    gf.LinkFile(gf.FILE_ILDG, pattern=timeslice).load(dims, out, back)
    assert np.array_equal(back, U)
    # the text records: up to the first NUL, xlf-info with the annealing steps
    recs = lime.parse(out.read_bytes())
    assert [r[0] for r in recs] == [r[0] for r in gold_template()]
    assert recs[0][3] == b"plaquette = 0.5876 beta = 6.0 SA steps 7"
    assert recs[3][3] == b"lfn://test/ensemble/cfg_0001"
    assert [(r[1], r[2]) for r in recs] == [(r[1], r[2]) for r in gold_template()]


@pytest.mark.parametrize("timeslice", [0, 1])
def test_qcdstag_writer_matches_reference_bytes(gf, tmp_path, timeslice):
    dims, U = tuple(GOLD["dims"]), gold_links()
    out = tmp_path / "out.qcdstag"
    gf.LinkFile(gf.FILE_QCDSTAG, pattern=timeslice).save(dims, out, U)
    assert sha(out) == GOLD["sha256"]["qcdstag_timeslice" if timeslice else "qcdstag"]


def test_vogt_writer_matches_reference_bytes(gf, tmp_path):
    dims, U = tuple(GOLD["dims"]), gold_links()
    out = tmp_path / "out.vogt"
    gf.LinkFile(gf.FILE_VOGT).save(dims, out, U)  # no header loaded: {4, 3, dims, 8}
    assert sha(out) == GOLD["sha256"]["vogt"]
    assert out.read_bytes()[:14] == struct.pack("<7h", 4, 3, *dims, 8)


# ------------------------------------------------------------ against the reference's code
DIMS4 = (2, 4, 6, 8)


def _links(dims, seed=7):
    return np.random.default_rng(seed).standard_normal(72 * int(np.prod(dims)))


@pytest.mark.parametrize("timeslice", [0, 1])
def test_standard_order_matches_reference_patterns(gf, refio, tmp_path, timeslice):
    dims = DIMS4
    V = int(np.prod(dims))
    U = np.arange(72 * V, dtype=np.float64)
    out = tmp_path / "plain"
    gf.LinkFile(gf.FILE_PLAIN, pattern=timeslice).save(dims, out, U)
    raw = np.fromfile(out, dtype=np.float64)
    want = np.array([refio.index_by_unique(dims, u, timeslice) for u in range(72 * V)], dtype=np.float64)
    assert np.array_equal(raw, want)


@pytest.mark.parametrize("timeslice", [0, 1])
def test_ildg_both_ways_with_reference(gf, refio, tmp_path, timeslice):
    dims = DIMS4
    V = int(np.prod(dims))
    U = _links(dims)
    tmpl = tmp_path / "template.lime"
    recs = [("xlf-info", True, False, b"beta 5.9\0"),
            ("ildg-format", False, False, b"<ildgFormat/>\0"),
            ("ildg-binary-data", False, True, bytes(8 * 72 * V))]
    refio.lime_write(tmpl, recs)
    ref_out, our_out = tmp_path / "ref.ildg", tmp_path / "our.ildg"
    refio.write_ildg(dims, tmpl, ref_out, U, 12, timeslice)
    f = gf.LinkFile(gf.FILE_ILDG, pattern=timeslice)
    f.read_header(dims, tmpl)
    f.save(dims, our_out, U, sa_steps=12)
    assert our_out.read_bytes() == ref_out.read_bytes()
    assert np.array_equal(refio.read_ildg(dims, our_out, V, timeslice), U)
    back = np.zeros_like(U)
    gf.LinkFile(gf.FILE_ILDG, pattern=timeslice).load(dims, ref_out, back)
    assert np.array_equal(back, U)


def test_ildg_without_template_reads_with_c_lime_and_reference(gf, refio, tmp_path):
    dims = DIMS4
    V = int(np.prod(dims))
    U = _links(dims, 3)
    out = tmp_path / "fresh.ildg"
    gf.LinkFile(gf.FILE_ILDG).save(dims, out, U)
    listed = refio.lime_list(out)
    assert [(t, mb, me) for t, mb, me, _ in listed] == [("ildg-format", True, False), ("ildg-binary-data", False, True)]
    assert listed[1][3] == 8 * 72 * V
    xml = refio.lime_read(out, 0, 4096).decode()
    assert "<field>su3gauge</field>" in xml and "<lx>4</lx>" in xml and "<lz>8</lz>" in xml and "<lt>2</lt>" in xml
    assert np.array_equal(refio.read_ildg(dims, out, V), U)


@pytest.mark.parametrize("timeslice", [0, 1])
def test_qcdstag_both_ways_with_reference(gf, refio, tmp_path, timeslice):
    dims = DIMS4
    V = int(np.prod(dims))
    U = _links(dims, 5)
    ref_out, our_out = tmp_path / "ref.q", tmp_path / "our.q"
    refio.write_qcdstag(dims, ref_out, U, timeslice)
    gf.LinkFile(gf.FILE_QCDSTAG, pattern=timeslice).save(dims, our_out, U)
    assert our_out.read_bytes() == ref_out.read_bytes()
    assert np.array_equal(refio.read_qcdstag(dims, our_out, V, timeslice), U)
    back = np.zeros_like(U)
    gf.LinkFile(gf.FILE_QCDSTAG, pattern=timeslice).load(dims, ref_out, back)
    assert np.array_equal(back, U)


def test_vogt_headers_with_reference(gf, refio, tmp_path):
    dims = DIMS4
    U = _links(dims, 9)
    ref_out = tmp_path / "ref.vogt"
    # a header in another order than the lattice's (the reference never checks it)
    refio.write_vogt(dims, [3, 8, 6, 4, 2, 8], ref_out, U)
    f = gf.LinkFile(gf.FILE_VOGT)
    back = np.zeros_like(U)
    f.load(dims, ref_out, back)
    assert np.array_equal(back, U)
    our_out = tmp_path / "our.vogt"
    f.save(dims, our_out, U)  # the loaded header is written back
    assert our_out.read_bytes() == ref_out.read_bytes()
    assert refio.read_vogt_header(our_out) == [4, 3, 8, 6, 4, 2, 8]


# ------------------------------------------------------------ round trips and errors
@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("reinterpret", ["REAL_STANDARD", "REAL_DOUBLE", "REAL_FLOAT"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("timeslice", [0, 1])
def test_round_trip(gf, tmp_path, fmt, reinterpret, dtype, timeslice):
    dims = (4, 2, 2, 6)
    U = _links(dims, 11).astype(dtype)
    kind = getattr(gf, fmt)
    f = gf.LinkFile(kind, getattr(gf, reinterpret), timeslice)
    p = tmp_path / "cfg"
    if kind == gf.FILE_HEADERONLY:
        # a file with a header of its own, which save copies
        p.write_bytes(b"#header of 21 bytes\n\0" + b"\0" * 72 * int(np.prod(dims)) *
                      (8 if reinterpret == "REAL_DOUBLE" or (reinterpret == "REAL_STANDARD" and dtype == np.float64) else 4))
        f.read_header(dims, p, dtype)
    f.save(dims, p, U, sa_steps=3)
    back = np.zeros_like(U)
    f.load(dims, p, back)
    exact = not (reinterpret == "REAL_FLOAT" and kind in (gf.FILE_PLAIN, gf.FILE_HEADERONLY, gf.FILE_VOGT))
    if exact:
        assert np.array_equal(back, U)
    else:
        assert np.array_equal(back, U.astype(np.float32).astype(dtype))
    if kind == gf.FILE_HEADERONLY:
        assert p.read_bytes().startswith(b"#header of 21 bytes\n\0")


def test_errors(gf, tmp_path):
    dims = (2, 2, 2, 2)
    U = np.zeros(72 * 16)
    with pytest.raises(gf.GFError) as e:
        gf.LinkFile(gf.FILE_PLAIN).load(dims, tmp_path / "missing", U)
    assert e.value.status == gf._lib.GF_ERR_IO
    short = tmp_path / "short"
    short.write_bytes(b"\0" * 100)
    for kind in (gf.FILE_PLAIN, gf.FILE_HEADERONLY, gf.FILE_QCDSTAG, gf.FILE_ILDG, gf.FILE_VOGT):
        with pytest.raises(gf.GFError) as e:
            gf.LinkFile(kind).load(dims, short, U)
        assert e.value.status == gf._lib.GF_ERR_IO
    bad = tmp_path / "bad.vogt"
    bad.write_bytes(struct.pack("<7h", 4, 3, 2, 2, 2, 4, 8) + b"\0" * 8 * 72 * 32)
    with pytest.raises(gf.GFError) as e:
        gf.LinkFile(gf.FILE_VOGT).load(dims, bad, U)  # volume 32 against 16
    assert e.value.status == gf._lib.GF_ERR_BAD_SHAPE
    wrong_real = tmp_path / "wrong.vogt"
    wrong_real.write_bytes(struct.pack("<7h", 4, 3, 2, 2, 2, 2, 4) + b"\0" * 8 * 72 * 16)
    with pytest.raises(gf.GFError):
        gf.LinkFile(gf.FILE_VOGT).load(dims, wrong_real, U)
    noildg = tmp_path / "no.ildg"
    noildg.write_bytes(lime.message([("xlf-info", True, True, b"x\0")]))
    with pytest.raises(gf.GFError) as e:
        gf.LinkFile(gf.FILE_ILDG).load(dims, noildg, U)
    assert "ildg-binary-data" in str(e.value)
    with pytest.raises(gf.GFError):
        gf.LinkFile(99)
