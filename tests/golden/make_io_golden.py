"""Writes tests/golden/io_vectors.json: SHA-256 digests of configuration files
written by the reference's own code (ildg.cpp, qcdstag.cpp, c-lime, FileVogt
and LinkFile's standard order composed from its pattern classes), compiled by
`make -C oracle ref` into oracle/_ref/libref_io.so.  The links are a seeded
normal sample (the file code does not need SU(3) matrices); the ILDG template
holds text records that end in a NUL, as the reference's writer copies each
text record up to its first NUL.

Usage (from the repository root, with /root/reference present):
    make -C oracle ref && python tests/golden/make_io_golden.py
"""
import hashlib
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from ref_io import RefIO  # noqa: E402

DIMS = (2, 4, 2, 6)  # Nt, Nx, Ny, Nz: the x and z extents differ from t and from each other
SEED = 2024
STEPS = 7


def links(dims=DIMS, seed=SEED):
    V = int(np.prod(dims))
    return np.random.default_rng(seed).standard_normal(72 * V)


def template_records(dims=DIMS):
    V = int(np.prod(dims))
    return [
        ("xlf-info", True, False, b"plaquette = 0.5876 beta = 6.0\0"),
        ("ildg-format", False, False,
         b"<?xml version=\"1.0\"?><ildgFormat><field>su3gauge</field><precision>64</precision>"
         b"<lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt></ildgFormat>\0" % (dims[1], dims[2], dims[3], dims[0])),
        ("ildg-binary-data", False, False, bytes(8 * 72 * V)),
        ("ildg-data-lfn", False, True, b"lfn://test/ensemble/cfg_0001\0"),
    ]


def sha(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def main():
    ref = RefIO()
    U = links()
    out = {"dims": list(DIMS), "seed": SEED, "sa_steps": STEPS,
           "template": [[t, mb, me, d.hex() if t != "ildg-binary-data" else len(d)] for t, mb, me, d in
                        template_records()]}
    digests = {}
    with tempfile.TemporaryDirectory() as d:
        d = Path(d)
        ref.lime_write(d / "template.lime", template_records())
        digests["template"] = sha(d / "template.lime")
        for ts in (0, 1):
            sfx = "_timeslice" if ts else ""
            ref.write_ildg(DIMS, d / "template.lime", d / f"ildg{sfx}", U, STEPS, ts)
            digests["ildg" + sfx] = sha(d / f"ildg{sfx}")
            ref.write_qcdstag(DIMS, d / f"qcdstag{sfx}", U, ts)
            digests["qcdstag" + sfx] = sha(d / f"qcdstag{sfx}")
        ref.write_vogt(DIMS, [3, *DIMS, 8], d / "vogt", U)
        digests["vogt"] = sha(d / "vogt")
    out["sha256"] = digests
    (HERE / "io_vectors.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "io_vectors.json")


if __name__ == "__main__":
    main()
