"""The host-side C++ drop-in headers (include/culgt) without a GPU: the
SiteIndex orderings and neighbour table against the reference-header golden
vectors and closed forms, and every header compiling on its own."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include" / "culgt").glob("*.h*"))
GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "ref_vectors.npz")


@pytest.fixture(scope="module")
def dump_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("culgt") / "site_index_dump"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-D_X_=4", "-D_T_=4", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "culgt" / "site_index_dump.cpp"), "-o", str(exe)], check=True)
    return exe


def run_dump(exe, dims):
    out = subprocess.run([str(exe), *map(str, dims)], capture_output=True, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        name, what, *vals = line.split()
        res[(name, what)] = np.array(vals, dtype=np.int64)
    return res


def lex(c, dims):
    return ((c[..., 0] * dims[1] + c[..., 1]) * dims[2] + c[..., 2]) * dims[3] + c[..., 3]


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 4, 8), (6, 4, 8, 4), (2, 2, 2, 2)])
def test_site_index_orderings(dump_exe, dims):
    d = run_dump(dump_exe, dims)
    V = int(np.prod(dims))
    key = "x".join(map(str, dims))
    # FULL_SPLIT is the GpuPattern order: the reference's own neighbour table
    assert np.array_equal(d[("FULL_SPLIT", "nn")], GOLDEN[f"nn_{key}"].astype(np.int64))
    c = np.stack(np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"), axis=-1).reshape(-1, 4)
    n = lex(c, dims)
    V3 = V // dims[0]
    n3 = lex(np.concatenate([np.zeros((len(c), 1), int), c[:, 1:]], axis=1), dims)
    idx = {
        "NO_SPLIT": n,
        "FULL_SPLIT": np.where(c.sum(axis=1) % 2 == 0, n // 2, n // 2 + V // 2),
        "TIMESLICE_SPLIT": c[:, 0] * V3 + np.where(c[:, 1:].sum(axis=1) % 2 == 0, n3 // 2, n3 // 2 + V3 // 2),
    }
    for name, ix in idx.items():
        # coordinates of index i and back
        co = d[(name, "coords")].reshape(V, 5)
        assert np.array_equal(co[:, 4], np.arange(V)), name
        want = np.empty((V, 4), int)
        want[ix] = c
        assert np.array_equal(co[:, :4], want), name
        # neighbour table: backward / forward in every direction, periodic
        nn = d[(name, "nn")].reshape(8, V)
        for mu in range(4):
            for up, step in ((0, -1), (1, 1)):
                cc = c.copy()
                cc[:, mu] = (cc[:, mu] + step) % dims[mu]
                assert np.array_equal(nn[2 * mu + up][ix], ix[np.argsort(n)][lex(cc, dims)]), (name, mu, up)
        # index within the timeslice and back
        sl = d[(name, "slice")].reshape(V, 2)
        assert np.array_equal(sl[:, 1], np.arange(V)), name
        # setNeighbour walks the table: forward in z, then backward in t
        st = d[(name, "step")]
        cc = c.copy()
        cc[:, 3] = (cc[:, 3] + 1) % dims[3]
        cc[:, 0] = (cc[:, 0] - 1) % dims[0]
        assert np.array_equal(st[ix], ix[np.argsort(n)][lex(cc, dims)]), name


@pytest.mark.parametrize("header", HEADERS, ids=lambda p: p.name)
def test_header_compiles_alone(tmp_path, header):
    src = tmp_path / "one.cpp"
    src.write_text(f'#include "culgt/{header.name}"\nint main() {{ return 0; }}\n')
    for defs in (["-D_X_=8", "-D_T_=8"], ["-D_X_=8", "-D_T_=4", "-D_Y_=6", "-D_Z_=4", "-DDOUBLEPRECISION"]):
        subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", *defs, f"-I{ROOT / 'include'}",
                        str(src)], check=True)
