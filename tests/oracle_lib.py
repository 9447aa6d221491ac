"""ctypes access to the CPU oracle (oracle/liboracle_*.so) and to the
reference-header harness (oracle/_ref/libsu3gf_ref.so).

Test infrastructure: these are the checkers the gfx950 results are compared
against; nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
# SU3GF_ORACLE_DIR points the tests at another build of the oracle, e.g. a
# copy with a planted mistake, to see which pins catch it
ORACLE_DIR = Path(os.environ.get("SU3GF_ORACLE_DIR", ROOT / "oracle"))

LANDAU, COULOMB, MAG, U1XU1 = range(4)
OR, MICRO, SA, SR, RANDOM = range(5)


def _ensure_built(path: Path) -> None:
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "all"], check=True)


def dims_arr(dims):
    return (C.c_int * 4)(*[int(d) for d in dims])


class Oracle:
    """The C restatement of the reference sweep (Real = double or float)."""

    def __init__(self, kind: str = "f64"):
        path = ORACLE_DIR / f"liboracle_{kind}.so"
        _ensure_built(path)
        self.lib = C.CDLL(str(path))
        self.kind = kind
        self.real = C.c_double if kind == "f64" else C.c_float
        self.np_real = np.float64 if kind == "f64" else np.float32
        L = self.lib
        P = C.POINTER(self.real)
        pd = C.POINTER(C.c_double)
        L.orc_volume.restype = C.c_int64
        L.orc_site_index.restype = C.c_int64
        L.orc_site_index.argtypes = [C.c_int * 4, C.c_int * 4]
        L.orc_site_coords.argtypes = [C.c_int * 4, C.c_int64, C.c_int * 4]
        L.orc_neighbour_table.argtypes = [C.c_int * 4, C.c_void_p]
        L.orc_pattern_index.restype = C.c_int64
        L.orc_pattern_index.argtypes = [C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_philox4x32_10.argtypes = [C.c_uint32 * 4, C.c_uint32 * 2, C.c_uint32 * 4]
        L.orc_u01_64.restype = C.c_double
        L.orc_u01_64.argtypes = [C.c_uint64]
        L.orc_u01_32.restype = C.c_float
        L.orc_u01_32.argtypes = [C.c_uint32]
        L.orc_philox_stream.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_int, P]
        L.orc_rng_tid.argtypes = [C.c_int64, C.c_int]
        for n in ("orc_reconstruct_third_line", "orc_project_su3"):
            getattr(L, n).argtypes = [P]
        L.orc_subgroup_quaternion.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_left_subgroup_mult.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_right_subgroup_mult.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_sweep.argtypes = [C.c_int, C.c_int, C.c_int * 4, P, C.c_int, C.c_float, C.c_int, C.c_uint32, C.c_int]
        L.orc_sweep_coulomb.argtypes = [C.c_int, C.c_int * 4, P, P, C.c_int, C.c_float, C.c_int, C.c_uint32, C.c_int]
        L.orc_project_all.argtypes = [C.c_int * 4, P]
        L.orc_set_hot.argtypes = [C.c_int * 4, P, C.c_int, C.c_uint32]
        L.orc_quality_per_site.argtypes = [C.c_int, C.c_int * 4, P, pd, pd]
        L.orc_quality.argtypes = [C.c_int, C.c_int * 4, P, pd, pd]
        L.orc_run_or.argtypes = [C.c_int, C.c_int * 4, P, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, pd, pd]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_update.argtypes = [C.c_int, P, C.c_float, C.c_int, C.c_int, C.c_uint32]

    def _p(self, arr: np.ndarray):
        assert arr.dtype == self.np_real and arr.flags.c_contiguous
        return arr.ctypes.data_as(C.POINTER(self.real))

    # geometry
    def neighbour_table(self, dims) -> np.ndarray:
        V = int(np.prod(dims))
        nn = np.zeros(8 * V, dtype=np.int32)
        self.lib.orc_neighbour_table(dims_arr(dims), nn.ctypes.data_as(C.c_void_p))
        return nn

    def site_index(self, dims, coord) -> int:
        return self.lib.orc_site_index(dims_arr(dims), dims_arr(coord))

    def site_coords(self, dims, idx):
        c = (C.c_int * 4)()
        self.lib.orc_site_coords(dims_arr(dims), idx, c)
        return tuple(c)

    # rng
    def philox(self, ctr, key):
        out = (C.c_uint32 * 4)()
        self.lib.orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
        return tuple(out)

    def philox_stream(self, tid, seed, counter, n):
        out = np.zeros(n, dtype=self.np_real)
        self.lib.orc_philox_stream(tid, seed, counter, n, self._p(out))
        return out

    # matrices (18 reals, row-major re/im)
    def mat_op(self, name, m, *args):
        m = np.array(m, dtype=self.np_real)
        getattr(self.lib, name)(self._p(m), *args)
        return m

    def subgroup_quaternion(self, m, i, j):
        m = np.array(m, dtype=self.np_real)
        q = np.zeros(4, dtype=self.np_real)
        self.lib.orc_subgroup_quaternion(self._p(m), i, j, self._p(q))
        return q

    def subgroup_mult(self, side, m, i, j, q):
        m = np.array(m, dtype=self.np_real)
        q = np.array(q, dtype=self.np_real)
        fn = self.lib.orc_left_subgroup_mult if side == "left" else self.lib.orc_right_subgroup_mult
        fn(self._p(m), i, j, self._p(q))
        return m

    # lattice operations (GpuPattern 18-real arrays, modified in place)
    def sweep(self, gauge, algo, dims, U, parity, param=0.0, seed=0, counter=0, nsb=32):
        rc = self.lib.orc_sweep(gauge, algo, dims_arr(dims), self._p(U), parity, C.c_float(param), seed,
                                C.c_uint32(counter & 0xFFFFFFFF), nsb)
        assert rc == 0

    def sweep_coulomb(self, algo, dims3, UtUp, UtDw, parity, param=0.0, seed=0, counter=0, nsb=32):
        rc = self.lib.orc_sweep_coulomb(algo, dims_arr(dims3), self._p(UtUp), self._p(UtDw), parity,
                                        C.c_float(param), seed, C.c_uint32(counter & 0xFFFFFFFF), nsb)
        assert rc == 0

    def project_all(self, dims, U):
        self.lib.orc_project_all(dims_arr(dims), self._p(U))

    def set_hot(self, dims, U, seed, counter):
        self.lib.orc_set_hot(dims_arr(dims), self._p(U), seed, C.c_uint32(counter & 0xFFFFFFFF))

    def quality(self, gauge, dims, U):
        g, t = C.c_double(), C.c_double()
        assert self.lib.orc_quality(gauge, dims_arr(dims), self._p(U), C.byref(g), C.byref(t)) == 0
        return g.value, t.value

    def quality_per_site(self, gauge, dims, U):
        V = int(np.prod(dims))
        g = np.zeros(V)
        a = np.zeros(V)
        pd = C.POINTER(C.c_double)
        assert self.lib.orc_quality_per_site(gauge, dims_arr(dims), self._p(U), g.ctypes.data_as(pd),
                                             a.ctypes.data_as(pd)) == 0
        return g, a

    def run_or(self, gauge, dims, U, omega, precision, maxiter, checkprecision, reproject):
        g, t = C.c_double(), C.c_double()
        n = self.lib.orc_run_or(gauge, dims_arr(dims), self._p(U), C.c_float(omega), C.c_float(precision), maxiter,
                                checkprecision, reproject, C.byref(g), C.byref(t))
        return n, g.value, t.value

    def update(self, algo, a, param=0.0, tid=0, seed=0, counter=0):
        """One update rule applied to the quaternion a (returns the new one)."""
        a = np.array(a, dtype=self.np_real)
        self.lib.orc_update(algo, self._p(a), C.c_float(param), tid, seed, C.c_uint32(counter & 0xFFFFFFFF))
        return a

    def set_threads(self, n: int) -> None:
        self.lib.orc_set_threads(n)


class RefHarness:
    """The reference's own lattice / Random123 templates compiled by
    oracle/Makefile (target ``ref``) into oracle/_ref/libsu3gf_ref.so."""

    PATH = ROOT / "oracle" / "_ref" / "libsu3gf_ref.so"

    @classmethod
    def available(cls) -> bool:
        return cls.PATH.exists()

    def __init__(self):
        self.lib = C.CDLL(str(self.PATH))
        L = self.lib
        P = C.POINTER(C.c_double)
        pd = C.POINTER(C.c_double)
        L.ref_neighbour_table.argtypes = [C.c_int * 4, C.c_void_p]
        L.ref_site_index.argtypes = [C.c_int * 4, C.c_int * 4]
        L.ref_site_coords.argtypes = [C.c_int * 4, C.c_int32, C.c_int * 4]
        L.ref_pattern_index.argtypes = [C.c_int * 4, C.c_int32, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_philox4x32.argtypes = [C.c_uint32 * 4, C.c_uint32 * 2, C.c_uint32 * 4]
        L.ref_u01_64.restype = C.c_double
        L.ref_u01_64.argtypes = [C.c_uint64]
        L.ref_u01_32.restype = C.c_float
        L.ref_u01_32.argtypes = [C.c_uint32]
        for n in ("ref_reconstruct_third_line", "ref_project_su3"):
            getattr(L, n).argtypes = [P]
        L.ref_subgroup_quaternion.argtypes = [P, C.c_int, C.c_int, P]
        L.ref_left_subgroup_mult.argtypes = [P, C.c_int, C.c_int, P]
        L.ref_right_subgroup_mult.argtypes = [P, C.c_int, C.c_int, P]
        L.ref_det.argtypes = [P, P]
        L.ref_sweep_landau.argtypes = [C.c_int, C.c_int * 4, P, C.c_int, C.c_float]
        L.ref_project_all.argtypes = [C.c_int * 4, P]
        L.ref_quality_landau.argtypes = [C.c_int * 4, P, pd, pd]

    @staticmethod
    def _p(a):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def neighbour_table(self, dims):
        V = int(np.prod(dims))
        nn = np.zeros(8 * V, dtype=np.int32)
        self.lib.ref_neighbour_table(dims_arr(dims), nn.ctypes.data_as(C.c_void_p))
        return nn

    def site_index(self, dims, coord):
        return self.lib.ref_site_index(dims_arr(dims), dims_arr(coord))

    def site_coords(self, dims, idx):
        c = (C.c_int * 4)()
        self.lib.ref_site_coords(dims_arr(dims), idx, c)
        return tuple(c)

    def pattern_index(self, dims, idx, mu, i, j, c):
        return self.lib.ref_pattern_index(dims_arr(dims), idx, mu, i, j, c)

    def philox(self, ctr, key):
        out = (C.c_uint32 * 4)()
        self.lib.ref_philox4x32((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
        return tuple(out)

    def mat_op(self, name, m, *args):
        m = np.array(m, dtype=np.float64)
        getattr(self.lib, name)(self._p(m), *args)
        return m

    def subgroup_quaternion(self, m, i, j):
        m = np.array(m, dtype=np.float64)
        q = np.zeros(4)
        self.lib.ref_subgroup_quaternion(self._p(m), i, j, self._p(q))
        return q

    def subgroup_mult(self, side, m, i, j, q):
        m = np.array(m, dtype=np.float64)
        q = np.array(q, dtype=np.float64)
        fn = self.lib.ref_left_subgroup_mult if side == "left" else self.lib.ref_right_subgroup_mult
        fn(self._p(m), i, j, self._p(q))
        return m

    def det(self, m):
        m = np.array(m, dtype=np.float64)
        out = np.zeros(2)
        self.lib.ref_det(self._p(m), self._p(out))
        return complex(out[0], out[1])

    def sweep_landau(self, algo, dims, U, parity, param):
        assert self.lib.ref_sweep_landau(algo, dims_arr(dims), self._p(U), parity, C.c_float(param)) == 0

    def project_all(self, dims, U):
        self.lib.ref_project_all(dims_arr(dims), self._p(U))

    def quality_landau(self, dims, U):
        g, t = C.c_double(), C.c_double()
        self.lib.ref_quality_landau(dims_arr(dims), self._p(U), C.byref(g), C.byref(t))
        return g.value, t.value


# ------------------------------------------------------------ layout helpers
def to_matrices(U: np.ndarray, V: int) -> np.ndarray:
    """GpuPattern 18-real array -> complex array [mu, idx, 3, 3]."""
    a = U.reshape(4, 3, 3, 2, V)  # mu, i, j, c, idx
    m = a[:, :, :, 0, :] + 1j * a[:, :, :, 1, :]
    return np.moveaxis(m, -1, 1)  # mu, idx, i, j


def from_matrices(M: np.ndarray) -> np.ndarray:
    """complex [mu, idx, 3, 3] -> GpuPattern 18-real array."""
    V = M.shape[1]
    m = np.moveaxis(M, 1, -1)  # mu, i, j, idx
    out = np.empty((4, 3, 3, 2, V))
    out[:, :, :, 0, :] = m.real
    out[:, :, :, 1, :] = m.imag
    return out.reshape(-1).copy()


def rows01(U: np.ndarray, V: int) -> np.ndarray:
    """The 48 reals per site the sweep reads and writes (rows 0,1)."""
    return U.reshape(4, 18, V)[:, :12, :]
