"""The C ABI boundary without a GPU: the library loads, exports every function
declared in include/*.h, rejects bad arguments on the host with the
documented status codes, and the host-side pieces (configuration generator,
the division-by-constant used for site decoding) are correct."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(gf_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported(gf):
    lib = gf._lib.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/*.h but not exported"
    # and the Python binding knows every one of them
    assert set(names) == set(gf._lib.SIGNATURES)


def test_version(gf):
    assert b"gfx950" in gf._lib.load().gf_version()


def _lat(gf, dims, dtype=1, layout=18):
    return gf._lib.gf_lattice((C.c_int * 4)(*dims), dtype, layout)


@pytest.mark.parametrize("dims,status", [((4, 4, 4, 3), 1), ((4, 5, 4, 4), 1), ((3, 4, 4, 4), 1), ((0, 4, 4, 4), 1)])
def test_sweep_rejects_bad_shapes(gf, dims, status):
    lib = gf._lib.load()
    lat = _lat(gf, dims)
    rc = lib.gf_sweep(C.byref(lat), C.c_void_p(0x1000), None, 0, 0, 0, C.c_float(1.7), 0, 0, 32, None)
    assert rc == status
    assert "odd" in gf._lib.last_error() or "positive" in gf._lib.last_error()


def test_sweep_rejects_bad_arguments(gf):
    lib = gf._lib.load()
    lat = _lat(gf, (4, 4, 4, 4))
    good = C.c_void_p(0x1000)
    assert lib.gf_sweep(C.byref(lat), None, None, 0, 0, 0, C.c_float(1.7), 0, 0, 32, None) == 2
    assert lib.gf_sweep(C.byref(lat), good, None, 2, 0, 0, C.c_float(1.7), 0, 0, 32, None) == 2
    assert lib.gf_sweep(C.byref(lat), good, None, 0, 7, 0, C.c_float(1.7), 0, 0, 32, None) == 2
    assert lib.gf_sweep(C.byref(lat), good, None, 0, 0, 9, C.c_float(1.7), 0, 0, 32, None) == 2
    # Coulomb needs a timeslice and the slice t-1 buffer
    assert lib.gf_sweep(C.byref(lat), good, good, 0, 1, 0, C.c_float(1.7), 0, 0, 32, None) == 1
    ts = _lat(gf, (1, 4, 4, 4))
    assert lib.gf_sweep(C.byref(ts), good, None, 0, 1, 0, C.c_float(1.7), 0, 0, 32, None) == 2
    bad_dtype = _lat(gf, (4, 4, 4, 4), dtype=5)
    assert lib.gf_sweep(C.byref(bad_dtype), good, None, 0, 0, 0, C.c_float(1.7), 0, 0, 32, None) == 2
    bad_layout = _lat(gf, (4, 4, 4, 4), layout=16)
    assert lib.gf_sweep(C.byref(bad_layout), good, None, 0, 0, 0, C.c_float(1.7), 0, 0, 32, None) == 2


def test_gauge_sweep_volume_limit(gf):
    """The gauge-transformation sweep's fields use 32-bit byte offsets: its
    drivers refuse more than 44,739,200 sites on the host, with the link
    sweeps named as the way out, before touching the device."""
    lib = gf._lib.load()
    big = _lat(gf, (84, 84, 84, 84))  # 49,787,136 sites
    good = C.c_void_p(0x1000)
    out = gf._lib.gf_run_result()
    for mode in (1, 2, 3):  # GF_MIXED_GAUGE32, GAUGE64, GAUGE24
        rc = lib.gf_run_mixed(C.byref(big), good, 0, 1.7, 1e-6, 1e-12, 10, mode, 100, 10, 100, C.byref(out), None)
        assert rc == 1, mode
        assert "44739200" in gf._lib.last_error() and "49787136" in gf._lib.last_error()
    ctr, temps = C.c_uint32(0), C.c_int(0)
    for mode in (1, 2):
        rc = lib.gf_run_sa(C.byref(big), good, None, 0, 1.4, 0.01, 10, 0, 100, mode, 1, C.byref(ctr), C.byref(temps),
                           None)
        assert rc == 1, mode
    big32 = _lat(gf, (84, 84, 84, 84), 0, 12)
    ms = C.c_float()
    for variant in (2, 3, 4):
        assert lib.gf_time_sweeps(C.byref(big32), good, 0, 10, variant, C.byref(ms), None) == 1, variant
    # the FP64 gauge-transformation sweep addresses 12 doubles per site with
    # 32-bit offsets: at most 22,369,600 sites, an FP64 lattice only
    big64 = _lat(gf, (70, 70, 70, 70))  # 24,010,000 sites
    assert lib.gf_time_sweeps(C.byref(big64), good, 0, 10, 7, C.byref(ms), None) == 1
    assert "22,369,600" in gf._lib.last_error()
    assert lib.gf_time_sweeps(C.byref(_lat(gf, (8, 8, 8, 8), 0, 12)), good, 0, 10, 7, C.byref(ms), None) == 2
    assert lib.gf_time_sweeps(C.byref(_lat(gf, (8, 8, 8, 8))), good, 0, 10, 8, C.byref(ms), None) == 2
    # modes past GF_MIXED_FP64 are unknown
    small = _lat(gf, (8, 8, 8, 8))
    assert lib.gf_run_mixed(C.byref(small), good, 0, 1.7, 1e-6, 1e-12, 10, 5, 100, 10, 100, C.byref(out), None) == 2
    assert "unknown mode" in gf._lib.last_error()


def test_batch_calls_check_every_problem_on_the_host(gf):
    """The batch drivers reject a bad batch before any device call, naming
    the problem; an empty batch is a no-op."""
    lib = gf._lib.load()
    good = _lat(gf, (4, 4, 4, 4))
    out = (gf._lib.gf_run_result * 2)()
    ptrs = (C.c_void_p * 2)(0x1000, 0x2000)
    lats = (gf._lib.gf_lattice * 2)(good, _lat(gf, (4, 4, 3, 4)))
    rc = lib.gf_batch_run_mixed(2, lats, ptrs, 0, 1.7, 1e-6, 1e-12, 10, 3, 50, 10, 100, 4, None, out)
    assert rc == 1 and "batch problem 1" in gf._lib.last_error()
    lats_ok = (gf._lib.gf_lattice * 2)(good, good)
    nullp = (C.c_void_p * 2)(0x1000, None)
    assert lib.gf_batch_run_mixed(2, lats_ok, nullp, 0, 1.7, 1e-6, 1e-12, 10, 3, 50, 10, 100, 4, None, out) == 2
    assert "batch problem 1" in gf._lib.last_error()
    assert lib.gf_batch_run_mixed(-1, lats_ok, ptrs, 0, 1.7, 1e-6, 1e-12, 10, 3, 50, 10, 100, 4, None, out) == 2
    assert lib.gf_batch_run_mixed(2, None, ptrs, 0, 1.7, 1e-6, 1e-12, 10, 3, 50, 10, 100, 4, None, out) == 2
    assert lib.gf_batch_run_mixed(0, None, None, 0, 1.7, 1e-6, 1e-12, 10, 3, 50, 10, 100, 4, None, None) == 0
    seeds = (C.c_int * 2)(1, 2)
    ctr = (C.c_uint32 * 2)(0, 0)
    assert lib.gf_batch_run_sa(2, lats, ptrs, 0, 1.0, 0.1, 10, 3, 100, 1, seeds, ctr, None, 4, None) == 1
    assert lib.gf_batch_run_sa(2, lats_ok, ptrs, 0, 1.0, 0.1, 10, 3, 100, 1, None, ctr, None, 4, None) == 2
    assert lib.gf_batch_run_sa(0, None, None, 0, 1.0, 0.1, 10, 3, 100, 1, None, None, None, 4, None) == 0
    # GF_SA_ONE_LAUNCH: modes 1 and 2 on FP64 lattices, Landau, Coulomb on a
    # 4D lattice and U1xU1, checked before anything runs
    one = gf.SA_ONE_LAUNCH
    assert lib.gf_batch_run_sa(2, lats_ok, ptrs, 0, 1.0, 0.1, 10, 3, 100, one | 0, seeds, ctr, None, 4, None) == 2
    assert "modes 1 and 2" in gf._lib.last_error()
    assert lib.gf_batch_run_sa(2, lats_ok, ptrs, gf.MAG, 1.0, 0.1, 10, 3, 100, one | 2, seeds, ctr, None, 4,
                               None) == 2
    f32 = gf.Lattice((4, 4, 4, 4), dtype=gf.F32).struct
    assert lib.gf_batch_run_sa(2, (gf._lib.gf_lattice * 2)(good, f32), ptrs, 0, 1.0, 0.1, 10, 3, 100, one | 2, seeds,
                               ctr, None, 4, None) == 2
    assert "batch problem 1" in gf._lib.last_error()
    slice3 = _lat(gf, (1, 4, 4, 4))
    assert lib.gf_batch_run_sa(2, (gf._lib.gf_lattice * 2)(slice3, good), ptrs, gf.COULOMB, 1.0, 0.1, 10, 3, 100,
                               one | 2, seeds, ctr, None, 4, None) == 2
    assert "batch problem 0" in gf._lib.last_error()
    assert lib.gf_batch_run_sa(2, lats, ptrs, 0, 1.0, 0.1, 10, 3, 100, one | 2, seeds, ctr, None, 4, None) == 1
    assert list(ctr) == [0, 0]
    prev = C.c_int()
    assert lib.gf_set_batch_launch_blocks(-1, C.byref(prev)) == 2
    assert lib.gf_set_batch_launch_blocks(4096, C.byref(prev)) == 0 and prev.value == 8192
    assert gf.set_batch_launch_blocks(prev.value) == 4096
    # the caller's streams: one per concurrent problem, none of them null
    short = (C.c_void_p * 4)(0x10, 0x20, None, 0x40)
    assert lib.gf_batch_run_mixed(2, lats_ok, ptrs, 0, 1.7, 1e-6, 1e-12, 10, 3, 50, 10, 100, 4, short, out) == 2
    # the Python mirror: one seed and one counter per problem, one stream per
    # concurrent problem, checked before the library is called
    two = [gf.Lattice((4, 4, 4, 4))] * 2
    with pytest.raises(ValueError):
        gf.batch_run_sa(two, [0x1000, 0x2000], seeds=[1], counters=[0, 0])
    with pytest.raises(ValueError):
        gf.batch_run_sa(two, [0x1000, 0x2000], seeds=[1, 2], counters=[0, 0, 0])
    with pytest.raises(ValueError):
        gf.batch_run_mixed(two, [0x1000, 0x2000], max_concurrent=4, streams=[0x10, 0x20])


def test_stats_rejects_bad_arguments(gf):
    lib = gf._lib.load()
    h = C.c_void_p()
    assert lib.gf_stats_create((C.c_int * 4)(4, 4, 4, 4), 9, 1, C.byref(h)) == 2
    assert lib.gf_stats_create((C.c_int * 4)(4, 4, 4, 4), 0, 5, C.byref(h)) == 2
    assert lib.gf_stats_create((C.c_int * 4)(4, 4, 3, 4), 0, 1, C.byref(h)) == 1


def test_lattice_bytes(gf):
    lib = gf._lib.load()
    assert lib.gf_lattice_bytes(C.byref(_lat(gf, (4, 4, 4, 4)))) == 256 * 4 * 18 * 8
    assert lib.gf_lattice_bytes(C.byref(_lat(gf, (4, 4, 4, 4), 0, 12))) == 256 * 4 * 12 * 4
    assert lib.gf_lattice_bytes(C.byref(_lat(gf, (4, 4, 4, 3)))) == 0


def test_python_mirror_checks_launch_geometry(gf):
    """The reference's launch (V/2/NSB blocks of 8*NSB threads) is validated
    before any device call (LandauGaugeFixingSU3_4D.cu:124-125)."""
    lat = gf.Lattice((4, 4, 4, 4))
    K = gf.LandauKernelsSU3(lat)
    with pytest.raises(gf.GFError):
        K.orStep(3, 256, 0x1000, None, 0, 1.7)
    with pytest.raises(gf.GFError):
        K.orStep(4, 250, 0x1000, None, 0, 1.7)
    with pytest.raises(gf.GFError):
        gf.CommonKernelsSU3(lat).projectSU3(4, 32, 0x1000)


def test_make_config_properties(gf):
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    U = gf.make_config(dims, 0.3, 1, 2)
    M = U.reshape(4, 3, 3, 2, V)
    M = np.moveaxis(M[..., 0, :] + 1j * M[..., 1, :], -1, 1)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    assert np.array_equal(U, gf.make_config(dims, 0.3, 1, 2))
    assert not np.array_equal(U, gf.make_config(dims, 0.3, 1, 3))
    assert not np.array_equal(U, gf.make_config(dims, 0.3, 2, 2))
    # sigma = 0 without a transformation: unit links
    C0 = gf.make_config(dims, 0.0, 1, 0).reshape(4, 3, 3, 2, V)
    assert np.array_equal(C0[:, :, :, 0, :], np.broadcast_to(np.eye(3)[None, :, :, None], (4, 3, 3, V)))
    with pytest.raises(gf.GFError):
        gf.make_config((4, 4, 4, 3), 0.3, 1, 2)


def test_make_config_link_distribution(gf):
    """U = exp(i H) with H = sum_a c_a lambda_a/2: Re tr U / 3 averages to the
    Gaussian expectation, close to 1 - sigma^2 * 4/3 * 1/2 for small sigma."""
    dims = (4, 4, 4, 8)
    sigma = 0.1
    U = gf.make_config(dims, sigma, 3, 0)
    V = int(np.prod(dims))
    A = U.reshape(4, 3, 3, 2, V)
    retr = (A[:, 0, 0, 0] + A[:, 1, 1, 0] + A[:, 2, 2, 0]).mean() / 3
    # <Re tr exp(iH)>/3 = 1 - <tr H^2>/6 + O(sigma^4), <tr H^2> = 8 sigma^2 / 2
    assert retr == pytest.approx(1 - 8 * sigma**2 / 12, abs=2e-3)


def _fastdiv(d):
    s = 0
    while (1 << s) < d:
        s += 1
    m = ((1 << 32) * ((1 << s) - d)) // d + 1
    return m & 0xFFFFFFFF, s


@pytest.mark.parametrize("d", [1, 2, 3, 4, 6, 8, 10, 12, 14, 16, 20, 24, 30, 32, 48, 63, 64, 96, 100, 128, 172, 256,
                               1000, 4096, 65537, 2**30 + 3, 2**31 - 1])
def test_fastdiv_formula(d):
    """q = (umulhi(n, m) + n) >> s, as the kernels decode site indices and
    map blocks to timeslices, for n < 2^31 (the volume limit checked by the
    library; block indices stay below it too)."""
    m, s = _fastdiv(d)
    rng = np.random.default_rng(d)
    ns = np.concatenate([np.arange(0, 5000), rng.integers(0, 2**31, 20000), [2**31 - 1, 2**31 - 2]]).astype(np.uint64)
    hi = (ns * np.uint64(m)) >> np.uint64(32)
    q = (hi + ns) >> np.uint64(s)
    assert np.array_equal(q, ns // np.uint64(d))


def test_sweep_path_settings(gf):
    """The tuning settings of the gauge-transformation sweep are host-side:
    gf_set_wavefront, gf_set_block_order and gf_set_pipelined_lag take -1
    (tuned) or a fixed value, return the previous one and reject the rest;
    gf_pipelined_launches counts the pipelined sweeps (none without a GPU)."""
    lib = gf._lib.load()
    prev = C.c_int()
    for name, good in (("gf_set_wavefront", [gf.WAVEFRONT_PIPELINED, 3, 0]), ("gf_set_block_order", [2, 0]),
                       ("gf_set_pipelined_lag", [256, 0])):
        fn = getattr(lib, name)
        assert fn(-1, C.byref(prev)) == 0
        first = prev.value
        for v in good:
            assert fn(v, None) == 0
            assert fn(v, C.byref(prev)) == 0 and prev.value == v
        assert fn(-2, None) == 2
        assert fn(first, None) == 0
    n = C.c_longlong(-1)
    assert lib.gf_pipelined_launches(C.byref(n)) == 0 and n.value == 0
    assert lib.gf_pipelined_launches(None) == 2
