"""gf_run_fast_noalloc, the OR driver that allocates nothing: its run state
lives in the library's fixed device array instead of memory the call
allocates.  Through the C ABI it returns what gf_run_fast returns -- sweeps,
checks, theta, gff and the links, bit for bit -- on two calls in a row, the
second on a smaller lattice (the array keeps the first call's contents), and
it rejects a lattice beyond its fixed scratch on the host."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import COULOMB, LANDAU, MAG


def _call(gf, name, lat, buf, gauge, prec, maxiter):
    r = gf._lib.gf_run_result()
    gf._lib.call(name, C.byref(lat.struct), C.c_void_p(buf.ptr), gauge, C.c_float(1.7), C.c_float(prec), maxiter,
                 100, C.byref(r), None)
    return r


FIELDS = ("sweeps", "sweeps_fp32", "checks", "gff", "theta", "converged")


@pytest.mark.gpu
@pytest.mark.parametrize("gauge,prec", [(LANDAU, 1e-12), (COULOMB, 1e-12), (MAG, 1e-10)])
def test_noalloc_matches_run_fast(gf, gpu, gauge, prec):
    for dims in [(8, 8, 8, 16), (4, 6, 4, 8)]:  # the second call's lattice is smaller
        print("noalloc", gauge, dims, flush=True)
        U = gf.make_config(dims, 0.3, 1, 2)
        lat = gf.Lattice(dims)
        d1 = gf.DeviceBuffer.from_numpy(U)
        d2 = gf.DeviceBuffer.from_numpy(U)
        r1 = _call(gf, "gf_run_fast", lat, d1, gauge, prec, 20000)
        r2 = _call(gf, "gf_run_fast_noalloc", lat, d2, gauge, prec, 20000)
        assert r2.converged == 1
        assert [getattr(r2, f) for f in FIELDS] == [getattr(r1, f) for f in FIELDS], dims
        assert np.array_equal(d2.download(np.float64, U.size), d1.download(np.float64, U.size)), dims


@pytest.mark.gpu
def test_noalloc_unconverged_and_fp32(gf, gpu):
    """A run stopped at maxiter (theta of the final state) and an FP32 run on
    the compact layout, against gf_run_fast."""
    dims = (4, 6, 4, 8)
    U = gf.make_config(dims, 0.3, 1, 2)
    lat = gf.Lattice(dims)
    d1, d2 = gf.DeviceBuffer.from_numpy(U), gf.DeviceBuffer.from_numpy(U)
    r1 = _call(gf, "gf_run_fast", lat, d1, COULOMB, 1e-12, 30)
    r2 = _call(gf, "gf_run_fast_noalloc", lat, d2, COULOMB, 1e-12, 30)
    assert r2.converged == 0 and r2.sweeps == 30
    assert [getattr(r2, f) for f in FIELDS] == [getattr(r1, f) for f in FIELDS]
    assert np.array_equal(d2.download(np.float64, U.size), d1.download(np.float64, U.size))
    lat32 = gf.Lattice(dims, gf.F32, gf.R12)
    b64 = gf.DeviceBuffer.from_numpy(U)
    out = []
    for name in ("gf_run_fast", "gf_run_fast_noalloc"):
        b32 = gf.DeviceBuffer(lat32.nbytes)
        gf._lib.call("gf_convert", C.byref(lat.struct), C.c_void_p(b64.ptr), C.byref(lat32.struct),
                     C.c_void_p(b32.ptr), 0, None)
        r = _call(gf, name, lat32, b32, LANDAU, 1e-6, 5000)
        out.append(([getattr(r, f) for f in FIELDS], b32.download(np.float32, lat32.count)))
    assert out[0][0] == out[1][0] and out[1][0][-1] == 1
    assert np.array_equal(out[0][1], out[1][1])


def test_noalloc_rejects_lattice_beyond_scratch(gf):
    """66 x 64^3 sites need 135168 blocks of 64 per half-sweep, more than the
    fixed scratch holds: rejected on the host before any device work (the
    link pointer is never dereferenced)."""
    lib = gf._lib.load()
    lat = gf._lib.gf_lattice((C.c_int * 4)(66, 64, 64, 64), 1, 18)
    r = gf._lib.gf_run_result()
    rc = lib.gf_run_fast_noalloc(C.byref(lat), C.c_void_p(0x1000), 0, C.c_float(1.7), C.c_float(1e-12), 10, 100,
                                 C.byref(r), None)
    assert rc == gf._lib.GF_ERR_BAD_SHAPE
    assert "scratch" in gf._lib.last_error()
