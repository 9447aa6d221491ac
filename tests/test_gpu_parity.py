"""Parity of the gfx950 kernels with the CPU oracle, called through the C ABI
via the Python mirror of the reference's classes.

The operator API (gf_sweep, gf_project, gf_set_hot, gf_stats_quality) runs the
kernels of the exact translation unit: multiplies and adds round separately,
in the reference's order, and the eight link contributions of a site are
added in the reference's ladder order.  For the deterministic FP64 updates
(OR, microcanonical, SR) the GPU result is therefore required to be
bit-identical to the oracle.  Updates that call transcendental functions (SA:
log and cospi; the random transform: sinpi and cospi) use OCML on the GPU and
glibc-based restatements in the oracle, which agree to a few ulp; those
compare within 1e-10.  The Philox streams are identical.
"""
import numpy as np
import pytest

from oracle_lib import COULOMB, LANDAU, MAG, MICRO, OR, RANDOM, SA, SR, U1XU1, rows01

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4, 4), (4, 6, 4, 8), (6, 4, 8, 4)]
KERNELS = {LANDAU: "LandauKernelsSU3", MAG: "MAGKernelsSU3", U1XU1: "U1xU1KernelsSU3"}


def test_reference_golden_lattice(gf, gpu):
    """The drop-in kernels against the reference's own code with no oracle in
    between: tests/golden/ref_vectors.npz holds what the reference's
    SiteIndex / SU3 / Quaternion classes, composed into an FP64 Landau sweep
    and compiled from the reference sources (oracle/ref_harness.cpp), made of
    one lattice.  projectSU3, three OR sweeps (omega 1.7) and then a
    microcanonical sweep through the C ABI give the same bits; the functional
    and theta of the input and of the swept lattice agree to rounding (the
    GPU reduction adds in another order: 1e-12 relative, 1e-14 absolute)."""
    from pathlib import Path
    g = np.load(Path(__file__).resolve().parent / "golden" / "ref_vectors.npz")
    dims = tuple(int(x) for x in g["lat_dims"])
    V = int(np.prod(dims))
    lat = gf.Lattice(dims)
    a, b = V // 2 // 32, 256
    d = gf.DeviceBuffer.from_numpy(g["lat_in"].copy())
    st = gf.GaugeFixingStats(gf.LandauKernelsSU3, d, lat)
    st.generateGaugeQuality()
    assert (st.getCurrentGff(), st.getCurrentA()) == pytest.approx(tuple(g["lat_quality_in"]), rel=1e-12, abs=1e-14)
    P = gf.DeviceBuffer.from_numpy(g["lat_in"].copy())
    gf.CommonKernelsSU3(lat).projectSU3(V // 32, 32, P)
    assert np.array_equal(P.download(np.float64, 72 * V), g["lat_project"])
    K = gf.LandauKernelsSU3(lat)
    for _ in range(3):
        for parity in (0, 1):
            K.orStep(a, b, d, None, parity, 1.7)
    assert np.array_equal(d.download(np.float64, 72 * V), g["lat_or3"])
    st.generateGaugeQuality()
    assert (st.getCurrentGff(), st.getCurrentA()) == pytest.approx(tuple(g["lat_quality_or3"]), rel=1e-12, abs=1e-14)
    for parity in (0, 1):
        K.microStep(a, b, d, None, parity)
    assert np.array_equal(d.download(np.float64, 72 * V), g["lat_or3_micro1"])


def _config(gf, dims, sigma=0.3, seed=1, haar=2):
    return gf.make_config(dims, sigma, seed, haar)


def _run_parity(gf, oracle, gauge, algo, dims, nsweeps, param, seed=7, counter0=11):
    V = int(np.prod(dims))
    lat = gf.Lattice(dims)
    U0 = _config(gf, dims)
    Uo = U0.copy()
    d = gf.DeviceBuffer.from_numpy(U0)
    K = getattr(gf, KERNELS[gauge])(lat)
    a, b = V // 2 // 32, 256
    counter = counter0
    for _ in range(nsweeps):
        for parity in (0, 1):
            if algo == OR:
                K.orStep(a, b, d, None, parity, param)
            elif algo == MICRO:
                K.microStep(a, b, d, None, parity)
            elif algo == SA:
                K.saStep(a, b, d, None, parity, param, seed, counter)
            elif algo == SR:
                K.srStep(a, b, d, None, parity, param, seed, counter)
            else:
                K.randomTrafo(a, b, d, None, parity, seed, counter)
            oracle.sweep(gauge, algo, dims, Uo, parity, param, seed, counter)
            counter += 1
    Ug = d.download(np.float64, U0.size)
    return U0, Uo, Ug, V


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("gauge", [LANDAU, MAG, U1XU1])
@pytest.mark.parametrize("algo", [OR, MICRO])
def test_sweep_bit_exact(gf, gpu, oracle, gauge, algo, dims):
    """FP64 OR / microcanonical half-sweeps reproduce the oracle bit for bit."""
    U0, Uo, Ug, V = _run_parity(gf, oracle, gauge, algo, dims, nsweeps=3, param=1.7)
    diff = np.abs(rows01(Ug, V) - rows01(Uo, V)).max()
    assert np.array_equal(rows01(Ug, V), rows01(Uo, V)), f"rows 0,1 differ, max {diff:.3e}"
    # the sweep never writes the third row (LandauKernelsSU3.hxx:255)
    assert np.array_equal(Ug.reshape(4, 18, V)[:, 12:, :], U0.reshape(4, 18, V)[:, 12:, :])
    assert not np.array_equal(rows01(Ug, V), rows01(U0, V))


@pytest.mark.parametrize("dp", [False, True])
@pytest.mark.parametrize("gauge", [LANDAU, MAG, U1XU1])
@pytest.mark.parametrize("algo,param", [(OR, 1.7), (MICRO, 0.0), (SA, 0.4)])
def test_fp32_sweeps(gf, gpu, oracle32, oracle32dp, gauge, algo, param, dp):
    """FP32 links (the reference's default build), and FP32 links with the
    update rule in double (GF_UPDATE_DP, its DPUPDATES build): three half-sweep
    pairs against the oracle's float and DPUPDATES builds, bit for bit for OR
    and microcanonical steps; the heatbath's log / cospi / sinpi are the
    device's and the host's, equal to float rounding."""
    oracle = oracle32dp if dp else oracle32
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    lat = gf.Lattice(dims, gf.F32)
    U0 = _config(gf, dims).astype(np.float32)
    Uo = U0.copy()
    d = gf.DeviceBuffer.from_numpy(U0)
    K = getattr(gf, KERNELS[gauge])(lat, dp_updates=dp)
    a, b = V // 2 // 32, 256
    counter = 11
    for _ in range(3):
        for parity in (0, 1):
            if algo == OR:
                K.orStep(a, b, d, None, parity, param)
            elif algo == MICRO:
                K.microStep(a, b, d, None, parity)
            else:
                K.saStep(a, b, d, None, parity, param, 7, counter)
            oracle.sweep(gauge, algo, dims, Uo, parity, param, 7, counter)
            counter += 1
    Ug = d.download(np.float32, U0.size)
    if algo == SA:
        # U1xU1 keeps (q0, q3) of the heatbath's element and normalises them,
        # which magnifies a last-ulp difference of the float logarithm
        tol = 1e-4 if gauge == U1XU1 and not dp else 1e-5
        assert np.abs(rows01(Ug, V) - rows01(Uo, V)).max() < tol
    else:
        assert np.array_equal(rows01(Ug, V), rows01(Uo, V)), np.abs(rows01(Ug, V) - rows01(Uo, V)).max()
    assert not np.array_equal(rows01(Ug, V), rows01(U0, V))


def test_fp32_dp_update_differs_from_fp32_update(gf, gpu, oracle32):
    """The flag changes the arithmetic: FP32 OR sweeps with the update in
    double leave the links a few float ulps away from the float-update sweeps
    (which the float oracle reproduces bit for bit)."""
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    lat = gf.Lattice(dims, gf.F32)
    U0 = _config(gf, dims).astype(np.float32)
    out = {}
    for dp in (False, True):
        d = gf.DeviceBuffer.from_numpy(U0)
        K = gf.LandauKernelsSU3(lat, dp_updates=dp)
        for parity in (0, 1, 0, 1):
            K.orStep(V // 2 // 32, 256, d, None, parity, 1.7)
        out[dp] = rows01(d.download(np.float32, U0.size), V)
    Uo = U0.copy()
    for parity in (0, 1, 0, 1):
        oracle32.sweep(LANDAU, OR, dims, Uo, parity, 1.7)
    assert np.array_equal(out[False], rows01(Uo, V))
    diff = np.abs(out[True] - out[False]).max()
    assert 0 < diff < 1e-5, diff


def test_sr_bit_exact(gf, gpu, oracle):
    dims = (4, 6, 4, 8)
    U0, Uo, Ug, V = _run_parity(gf, oracle, MAG, SR, dims, nsweeps=3, param=0.5)
    assert np.array_equal(rows01(Ug, V), rows01(Uo, V))


@pytest.mark.parametrize("gauge", [LANDAU, U1XU1])
def test_micro_preserves_functional(gf, gpu, gauge):
    """A microcanonical step replaces the local functional's maximiser direction
    by its reflection, so the Landau functional is unchanged (exactly, in exact
    arithmetic) however many sweeps are done."""
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    lat = gf.Lattice(dims)
    d = gf.DeviceBuffer.from_numpy(_config(gf, dims))
    st = gf.GaugeFixingStats(gf.LandauKernelsSU3, d, lat)
    st.generateGaugeQuality()
    g0, t0 = st.getCurrentGff(), st.getCurrentA()
    K = getattr(gf, KERNELS[gauge])(lat)
    for _ in range(10):
        for parity in (0, 1):
            K.microStep(V // 2 // 32, 256, d, None, parity)
    st.generateGaugeQuality()
    assert st.getCurrentGff() == pytest.approx(g0, rel=1e-12, abs=1e-13)
    assert st.getCurrentA() != pytest.approx(t0, rel=1e-3)  # the links did move


@pytest.mark.parametrize("gauge", [LANDAU, MAG, U1XU1])
@pytest.mark.parametrize("algo,param", [(SA, 0.4), (RANDOM, 0.0)])
def test_sweep_transcendental_updates(gf, gpu, oracle, gauge, algo, param):
    dims = (4, 6, 4, 8)
    U0, Uo, Ug, V = _run_parity(gf, oracle, gauge, algo, dims, nsweeps=2, param=param)
    err = np.abs(rows01(Ug, V) - rows01(Uo, V)).max()
    assert err < 1e-10, f"rows 0,1 differ by {err}"


@pytest.mark.parametrize("algo,param", [(OR, 1.7), (MICRO, 0.0), (SA, 0.7), (RANDOM, 0.0)])
def test_coulomb_slice_sweep(gf, gpu, oracle, algo, param):
    dims3 = (1, 8, 6, 4)
    V = int(np.prod(dims3))
    lat = gf.Lattice(dims3)
    up0 = _config(gf, dims3, seed=3, haar=4)
    dw0 = _config(gf, dims3, seed=5, haar=6)
    upo, dwo = up0.copy(), dw0.copy()
    dup, ddw = gf.DeviceBuffer.from_numpy(up0), gf.DeviceBuffer.from_numpy(dw0)
    K = gf.CoulombKernelsSU3(lat)
    a, b = V // 2 // 32, 256
    for it in range(3):
        for parity in (0, 1):
            c = 100 + 2 * it + parity
            if algo == OR:
                K.orStep(a, b, dup, ddw, None, parity, param)
            elif algo == MICRO:
                K.microStep(a, b, dup, ddw, None, parity)
            elif algo == SA:
                K.saStep(a, b, dup, ddw, None, parity, param, 9, c)
            else:
                K.randomTrafo(a, b, dup, ddw, None, parity, 9, c)
            oracle.sweep_coulomb(algo, dims3, upo, dwo, parity, param, 9, c)
    upg, dwg = dup.download(np.float64, up0.size), ddw.download(np.float64, dw0.size)
    assert np.abs(rows01(upg, V) - rows01(upo, V)).max() < 1e-10
    assert np.abs(rows01(dwg, V) - rows01(dwo, V)).max() < 1e-10
    # only the time links of slice t-1 are touched
    assert np.array_equal(dwg.reshape(4, 18, V)[1:], dw0.reshape(4, 18, V)[1:])


@pytest.mark.parametrize("gauge", [LANDAU, MAG, U1XU1])
def test_quality(gf, gpu, oracle, gauge):
    dims = (4, 6, 4, 8)
    U = _config(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    lat = gf.Lattice(dims)
    K = getattr(gf, KERNELS[gauge])
    st = gf.GaugeFixingStats(K, d, lat)
    st.generateGaugeQuality()
    g_ref, t_ref = oracle.quality(gauge, dims, U)
    assert st.getCurrentGff() == pytest.approx(g_ref, rel=1e-13)
    assert st.getCurrentA() == pytest.approx(t_ref, rel=1e-12)
    # per-site values
    V = int(np.prod(dims))
    dg, da = gf.DeviceBuffer(8 * V), gf.DeviceBuffer(8 * V)
    K(lat).generateGaugeQualityPerSite(V // 32, 32, d, dg, da)
    go, ao = oracle.quality_per_site(gauge, dims, U)
    assert np.allclose(dg.download(), go, rtol=1e-12, atol=1e-13)
    assert np.allclose(da.download(), ao, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("gauge", [LANDAU, MAG])
def test_quality_block_order(gf, gpu, oracle, gauge):
    """The drop-in quality kernel on a lattice whose XCDs each take 2 and 3
    timeslices, in every XCD block order (one pass, passes of 1) and in the
    order its tuner picks: per-site values, the reduced gff and theta (AVERAGE
    and MAX) identical bit for bit across orders, and the oracle's."""
    import ctypes as C
    for dims in ((16, 4, 8, 8), (24, 4, 4, 16)):
        V = int(np.prod(dims))
        U = _config(gf, dims)
        d = gf.DeviceBuffer.from_numpy(U)
        lat = gf.Lattice(dims)
        K = getattr(gf, KERNELS[gauge])
        prev = C.c_int()
        gf._lib.call("gf_set_block_order", 0, C.byref(prev))
        got = []
        try:
            for tpass in (0, 1, -1):
                gf._lib.call("gf_set_block_order", tpass, None)
                st = gf.GaugeFixingStats(K, d, lat)
                st.generateGaugeQuality()
                st_max = gf.GaugeFixingStats(K, d, lat, crit=gf.CRIT_MAX)
                st_max.generateGaugeQuality()
                dg, da = gf.DeviceBuffer(8 * V), gf.DeviceBuffer(8 * V)
                K(lat).generateGaugeQualityPerSite(V // 32, 32, d, dg, da)
                got.append((st.getCurrentGff(), st.getCurrentA(), st_max.getCurrentA(), dg.download(), da.download()))
        finally:
            gf._lib.call("gf_set_block_order", prev.value, None)
        for g in got[1:]:
            assert g[:3] == got[0][:3], dims
            assert np.array_equal(g[3], got[0][3]) and np.array_equal(g[4], got[0][4]), dims
        g_ref, t_ref = oracle.quality(gauge, dims, U)
        assert got[0][0] == pytest.approx(g_ref, rel=1e-13)
        assert got[0][1] == pytest.approx(t_ref, rel=1e-12)


def test_quality_coulomb_and_max(gf, gpu, oracle):
    dims3 = (1, 8, 6, 4)
    U = _config(gf, dims3, seed=3, haar=4)
    d = gf.DeviceBuffer.from_numpy(U)
    lat = gf.Lattice(dims3)
    st = gf.GaugeFixingStats(gf.CoulombKernelsSU3, d, lat)
    st.generateGaugeQuality()
    g_ref, t_ref = oracle.quality(COULOMB, dims3, U)
    assert st.getCurrentGff() == pytest.approx(g_ref, rel=1e-13)
    assert st.getCurrentA() == pytest.approx(t_ref, rel=1e-12)
    # StoppingCrit MAX: theta = max_x theta_x / Nc
    st_max = gf.GaugeFixingStats(gf.CoulombKernelsSU3, d, lat, crit=gf.CRIT_MAX)
    st_max.generateGaugeQuality()
    _, ao = oracle.quality_per_site(COULOMB, dims3, U)
    assert st_max.getCurrentA() == pytest.approx(ao.max() / 3.0, rel=1e-12)


def test_project_and_set_hot(gf, gpu, oracle):
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    lat = gf.Lattice(dims)
    U = _config(gf, dims)
    # perturb rows 0,1 off the group so that projection has work to do
    U = U + 1e-3 * np.random.default_rng(0).standard_normal(U.size)
    d = gf.DeviceBuffer.from_numpy(U)
    gf.CommonKernelsSU3(lat).projectSU3(V // 32, 32, d)
    Uo = U.copy()
    oracle.project_all(dims, Uo)
    assert np.abs(d.download() - Uo).max() < 1e-13
    h = gf.DeviceBuffer(lat.nbytes)
    gf.CommonKernelsSU3(lat).setHot(V // 32, 32, h, None, 5, 17)
    Uh = np.zeros(V * 72)
    oracle.set_hot(dims, Uh, 5, 17)
    assert np.abs(h.download() - Uh).max() < 1e-13


def test_convert_roundtrip(gf, gpu):
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    lat64 = gf.Lattice(dims)
    lat32 = gf.Lattice(dims, gf.F32, gf.R12)
    U = _config(gf, dims)
    d64 = gf.DeviceBuffer.from_numpy(U)
    d32 = gf.DeviceBuffer(lat32.nbytes)
    back = gf.DeviceBuffer(lat64.nbytes)
    import ctypes as C
    gf._lib.call("gf_convert", C.byref(lat64.struct), C.c_void_p(d64.ptr), C.byref(lat32.struct),
                 C.c_void_p(d32.ptr), 0, None)
    f = d32.download(np.float32, lat32.count).reshape(4, 12, V)
    assert np.array_equal(f, rows01(U, V).astype(np.float32))
    gf._lib.call("gf_convert", C.byref(lat32.struct), C.c_void_p(d32.ptr), C.byref(lat64.struct),
                 C.c_void_p(back.ptr), 1, None)
    Ub = back.download()
    assert np.abs(Ub - U).max() < 1e-6


def test_run_or_config0_matches_oracle(gf, gpu, oracle):
    """8^4 Landau, sigma 0.3 seed 1, Haar seed 2, omega 1.7, theta < 1e-12,
    checked every sweep, reprojected every 100 sweeps (including sweep 0),
    with the reference's loop (LandauGaugeFixingSU3_4D.cu:244-262)."""
    dims = (8, 8, 8, 8)
    U = _config(gf, dims, 0.3, 1, 2)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_or(gf.Lattice(dims), d, gf.LANDAU, 1.7, 1e-12, 5000, 1, 100)
    Uo = U.copy()
    n, g, t = oracle.run_or(LANDAU, dims, Uo, 1.7, 1e-12, 5000, 1, 100)
    assert res.converged and t < 1e-12
    assert res.sweeps == n
    assert res.gff == pytest.approx(g, rel=1e-13)
    assert np.array_equal(rows01(d.download(), int(np.prod(dims))), rows01(Uo, int(np.prod(dims))))


def test_run_fast_matches_oracle(gf, gpu, oracle):
    """Fused-check driver (contracted FP64): same stopping sweep (+-1) and the
    functional within 1e-10 of the oracle's reference-loop run."""
    dims = (8, 8, 8, 8)
    U = _config(gf, dims, 0.3, 1, 2)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_fast(gf.Lattice(dims), d, LANDAU, 1.7, 1e-12, 5000, 100)
    Uo = U.copy()
    n, g, t = oracle.run_or(LANDAU, dims, Uo, 1.7, 1e-12, 5000, 1, 100)
    assert res.converged
    assert abs(res.sweeps - n) <= 1
    assert res.gff == pytest.approx(g, rel=1e-10)
    assert res.theta < 1e-12


@pytest.mark.parametrize("gauge", [U1XU1, MAG])
def test_run_fast_fixed_sweeps(gf, gpu, oracle, gauge):
    """U1xU1 is checked with the Landau theta (U1xU1KernelsSU3.hxx:114-191), which
    the diagonal update never drives to zero, and MAG with its own theta; both
    run a fixed number of sweeps here and the functional is compared."""
    dims = (8, 8, 8, 8)
    U = _config(gf, dims, 0.3, 1, 2)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_fast(gf.Lattice(dims), d, gauge, 1.7, 0.0, 150, 100)
    Uo = U.copy()
    n, g, t = oracle.run_or(gauge, dims, Uo, 1.7, 0.0, 150, 1, 100)
    g_end, t_end = oracle.quality(gauge, dims, Uo)
    assert not res.converged and res.sweeps == n == 150
    assert res.gff == pytest.approx(g_end, rel=1e-10)
    assert res.theta == pytest.approx(t_end, rel=1e-6)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("dims,sigma,switch", [((8, 8, 8, 8), 0.3, 1e-6), ((12, 12, 12, 12), 0.25, 1e-6),
                                               ((8, 8, 8, 8), 0.3, 1e-12), ((4, 8, 8, 16), 0.3, 1e-6),
                                               ((16, 4, 4, 8), 0.3, 1e-6)])
def test_run_mixed_matches_fp64_oracle(gf, gpu, oracle, dims, sigma, switch, mode):
    """FP32 phase (links + g, or g alone in FP32 / FP64 arithmetic, or on
    24-bit fixed-point links), then FP64: the final functional is the FP64
    oracle's to 1e-10 (same gauge orbit, same maximum).  The shapes take each
    block order of the gauge-transformation sweep: 8^4 and 16 x 4 x 4 x 8 the
    timeslice-walking XCD order (Nt a multiple of 8, a timeslice's sites of a
    parity whole blocks; one and two timeslices per XCD), 4 x 8 x 8 x 16 the
    plain XCD order (32 blocks, Nt = 4), 12^4 the launch order (162 blocks)."""
    U = _config(gf, dims, sigma, 1, 2)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_mixed(gf.Lattice(dims), d, gf.LANDAU, 1.7, switch_precision=switch, precision=1e-12,
                       maxiter=20000, mode=mode, refresh=50, check32=10, reproject=100)
    Uo = U.copy()
    n, g, t = oracle.run_or(LANDAU, dims, Uo, 1.7, 1e-12, 20000, 1, 100)
    assert res.converged == 1 and res.theta < 1e-12
    assert 0 < res.sweeps_fp32 <= res.sweeps
    assert res.gff == pytest.approx(g, rel=1e-10)
    # the result is a gauge transform of the input: the plaquette is unchanged
    V = int(np.prod(dims))
    from oracle_lib import to_matrices
    assert _plaquette(to_matrices(d.download(), V), dims) == pytest.approx(_plaquette(to_matrices(U, V), dims), rel=1e-13)


def _plaquette(M, dims, every=1):
    """Average Re tr U_{mu nu}(x) / 3 over all planes and the sites whose
    lexicographic index is a multiple of `every` (gauge invariant plaquette
    by plaquette, so a sample of sites is an invariant too)."""
    V = int(np.prod(dims))
    # parity-split index of every sampled lexicographic site and its forward neighbours
    n = np.arange(0, V, every)
    c = np.stack(np.unravel_index(n, dims), axis=1)
    def pidx(cc):
        lex = np.ravel_multi_index(cc.T, dims)
        return np.where(cc.sum(axis=1) % 2 == 0, lex // 2, lex // 2 + V // 2)
    here = pidx(c)
    tot = 0.0
    for mu in range(4):
        cm = c.copy(); cm[:, mu] = (cm[:, mu] + 1) % dims[mu]
        xm = pidx(cm)
        for nu in range(mu + 1, 4):
            cn = c.copy(); cn[:, nu] = (cn[:, nu] + 1) % dims[nu]
            xn = pidx(cn)
            P = M[mu, here] @ M[nu, xm] @ np.conj(np.swapaxes(M[mu, xn], -1, -2)) @ np.conj(np.swapaxes(M[nu, here], -1, -2))
            tot += np.trace(P, axis1=-2, axis2=-1).real.sum()
    return tot / (6 * len(n) * 3)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_run_mixed_fp32_stagnation(gf, gpu, oracle, mode):
    """With a switch threshold FP32 cannot reach, the FP32 phase ends on
    stagnation and the FP64 phase still converges to the oracle's maximum."""
    dims = (8, 8, 8, 8)
    U = _config(gf, dims, 0.3, 1, 2)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_mixed(gf.Lattice(dims), d, gf.LANDAU, 1.7, switch_precision=1e-30, precision=1e-12,
                       maxiter=20000, mode=mode, refresh=50, check32=5, reproject=100)
    Uo = U.copy()
    n, g, t = oracle.run_or(LANDAU, dims, Uo, 1.7, 1e-12, 20000, 1, 100)
    assert res.converged == 1 and res.theta < 1e-12
    assert 0 < res.sweeps_fp32 < res.sweeps
    assert res.gff == pytest.approx(g, rel=1e-10)
