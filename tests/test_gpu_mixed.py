"""The single-precision gauge-transformation sweep and the Coulomb lattice
fixed all timeslices at once, against the FP64 oracle.

The gauge-transformation sweep keeps the links U0 (an FP32 copy) fixed and
updates g(x); the links it stands for are g(x) U0_mu(x) g(x+mu)^dagger, and
the drivers apply g to the FP64 configuration when the FP32 phase ends.  A
few such sweeps therefore track the oracle's FP64 link sweeps to FP32
accuracy, and a full mixed run lands on the oracle's maximum.
"""
import numpy as np
import pytest

from app_schedule import to_timeslices
from oracle_lib import COULOMB, LANDAU, MAG, OR, U1XU1, rows01

pytestmark = pytest.mark.gpu


def _cfg(gf, dims, sigma=0.3, seed=1, haar=2):
    return gf.make_config(dims, sigma, seed, haar)


@pytest.mark.parametrize("mode,tol", [(1, 2e-5), (2, 2e-6), (3, 2e-5)])
@pytest.mark.parametrize("gauge", [LANDAU, U1XU1, MAG])
def test_gauge_sweeps_track_fp64_link_sweeps(gf, gpu, oracle, mode, tol, gauge):
    """Three gauge-transformation sweeps (run_mixed with maxiter = 3 and a
    switch threshold never met), applied to U0 and projected, against three
    FP64 OR sweeps of the oracle followed by projectSU3."""
    # U1xU1 normalises (q0, q3) alone, which magnifies the FP32 storage
    # rounding of small rotations: a looser bound; MAG's update is quadratic
    # in the links and normalises (q0, q1, q2), likewise
    if gauge in (U1XU1, MAG):
        tol *= 5
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    U = _cfg(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_mixed(gf.Lattice(dims), d, gauge, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=3,
                       mode=mode, check32=10)
    assert res.sweeps == res.sweeps_fp32 == 3
    Uo = U.copy()
    for _ in range(3):
        for parity in (0, 1):
            oracle.sweep(gauge, OR, dims, Uo, parity, 1.7)
    oracle.project_all(dims, Uo)
    err = np.abs(rows01(d.download(), V) - rows01(Uo, V)).max()
    assert err < tol, err


@pytest.mark.parametrize("mode,tol", [(1, 2e-5), (2, 2e-6), (3, 2e-5)])
def test_coulomb_gauge_sweeps_track_fp64_slice_sweeps(gf, gpu, oracle, mode, tol):
    """Coulomb on a 4D lattice: three gauge-transformation sweeps of every
    timeslice at once (blocks of three waves, one per spatial direction),
    applied to U0 and projected, against three FP64 OR sweeps of each slice
    by the oracle followed by projectSU3, compared on every slice's spatial
    links.  The slices are independent (no time term in the Coulomb update);
    the oracle sweeps slice t in the 4D checkerboard's order, starting with
    the sites whose 3D parity is t mod 2."""
    dims = (4, 6, 4, 8)
    dims3 = (1,) + dims[1:]
    V3 = int(np.prod(dims[1:]))
    U = _cfg(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_mixed(gf.Lattice(dims), d, COULOMB, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=3,
                       mode=mode, check32=10)
    assert res.sweeps == res.sweeps_fp32 == 3
    S = to_timeslices(U, dims)
    got = to_timeslices(d.download(), dims)
    for t in range(dims[0]):
        up, dw = S[t].reshape(-1).copy(), S[t - 1].reshape(-1).copy()
        for _ in range(3):
            for parity in (t % 2, 1 - t % 2):
                oracle.sweep_coulomb(OR, dims3, up, dw, parity, 1.7)
        oracle.project_all(dims3, up)
        spatial = lambda a: a.reshape(4, 18, V3)[1:, :12, :]  # noqa: E731
        err = np.abs(spatial(got[t]) - spatial(up)).max()
        assert err < tol, (t, err)


@pytest.mark.parametrize("own_streams", [False, True])
def test_batch_runs_match_single_runs(gf, gpu, own_streams):
    """gf_batch_run_sa and then gf_batch_run_mixed on three copies of an 8^4
    configuration (Haar seeds 10-12), a 12^4 lattice and a 16 x 32^3 lattice
    (4096 blocks per half-sweep: it runs after the others, alone), on streams
    of the call's own or of the caller's: every problem's links, RNG counter,
    temperatures and run results are those of the single-lattice calls with
    the same arguments, bit for bit."""
    import ctypes as C

    streams = None
    if own_streams:
        streams = []
        for _ in range(4):
            s = C.c_void_p()
            gf._lib.call("gf_stream_create", C.byref(s))
            streams.append(s)
    shapes = [(8, 8, 8, 8)] * 3 + [(12, 12, 12, 12), (16, 32, 32, 32)]
    cfgs = [gf.make_config(d, 0.3, 1, 10 + k) for k, d in enumerate(shapes)]
    lats = [gf.Lattice(d) for d in shapes]
    seeds, counters = [3, 4, 5, 6, 7], [0, 1000, 2000, 3000, 4000]
    bufs = [gf.DeviceBuffer.from_numpy(c) for c in cfgs]
    sa = gf.batch_run_sa(lats, bufs, gf.LANDAU, 1.0, 0.05, 20, 3, 100, seeds, counters, mode=2, max_concurrent=4,
                         streams=streams)
    res = gf.batch_run_mixed(lats, bufs, gf.LANDAU, 1.7, 1e-8, 1e-11, 5000, check32=5, max_concurrent=4,
                             streams=streams)
    for k, dims in enumerate(shapes):
        d = gf.DeviceBuffer.from_numpy(cfgs[k])
        s1 = gf.run_sa(lats[k], d, gf.LANDAU, 1.0, 0.05, 20, 3, 100, seeds[k], counters[k], mode=2)
        r1 = gf.run_mixed(lats[k], d, gf.LANDAU, 1.7, 1e-8, 1e-11, 5000, check32=5)
        assert sa[k] == s1, (k, sa[k], s1)
        fields = ("sweeps", "sweeps_fp32", "checks", "gff", "theta", "converged")
        assert [getattr(res[k], f) for f in fields] == [getattr(r1, f) for f in fields], k
        assert res[k].converged == 1
        assert np.array_equal(bufs[k].download(np.float64, cfgs[k].size), d.download(np.float64, cfgs[k].size)), k
    for s in streams or []:
        gf._lib.call("gf_stream_destroy", s)


@pytest.mark.parametrize("gauge,mode,graph,limit", [(LANDAU, 2, False, 8192), (LANDAU, 1, False, 100),
                                                    (LANDAU, 2, True, 0), (COULOMB, 2, False, 8192),
                                                    (U1XU1, 1, False, 8192), (U1XU1, 2, True, 50)])
def test_one_launch_batch_annealing_matches_single_runs(gf, gpu, gauge, mode, graph, limit):
    """GF_SA_ONE_LAUNCH: lattices of several shapes (odd block counts, a
    last block with idle lanes, extents 4 to 12) annealed with each half-sweep
    of a group of them in one launch -- one group, several groups under a
    small block limit (a lattice above the limit alone), no limit, and the
    groups' schedules captured into graphs -- give every problem the links,
    RNG counter and temperature count of gf_run_sa on that lattice alone, bit
    for bit.  The problems have their own seeds and starting counters."""
    shapes = [(8, 8, 8, 8), (4, 6, 4, 8), (12, 12, 12, 12), (8, 8, 8, 8), (6, 6, 6, 10), (4, 4, 4, 4)]
    sigma = 0.3
    cfgs = [gf.make_config(d, sigma, 1 + k, 10 + k) for k, d in enumerate(shapes)]
    lats = [gf.Lattice(d) for d in shapes]
    seeds = [3, 4, 5, 3, 7, 8]
    counters = [0, 1000, 2000, 2**32 - 6, 3, 77]  # one counter wraps past 2^32
    kw = dict(tmax=2.8, tmin=1.9, steps=10) if gauge == U1XU1 else dict(tmax=1.45, tmin=0.01, steps=12)
    prev = gf.set_batch_launch_blocks(limit)
    try:
        bufs = [gf.DeviceBuffer.from_numpy(c) for c in cfgs]
        out = gf.batch_run_sa(lats, bufs, gauge, micro=2, reproject=5, seeds=seeds, counters=counters, mode=mode,
                              graph=graph, one_launch=True, **kw)
    finally:
        gf.set_batch_launch_blocks(prev)
    for k, dims in enumerate(shapes):
        d = gf.DeviceBuffer.from_numpy(cfgs[k])
        ref = gf.run_sa(lats[k], d, gauge, micro=2, reproject=5, seed=seeds[k], counter=counters[k], mode=mode, **kw)
        gf.synchronize()
        assert out[k] == ref, (k, out[k], ref)
        assert np.array_equal(bufs[k].download(np.float64, cfgs[k].size), d.download(np.float64, cfgs[k].size)), k


def test_check_count_when_fp32_phase_takes_every_sweep(gf, gpu):
    """A mixed run whose FP32 phase uses up maxiter leaves the FP64 phase no
    sweeps; the reported checks are the FP32 phase's alone.  With check32 = 5
    and theta far above the switch threshold the phase checks every 25 sweeps:
    after sweeps 25 and 50 of 60."""
    dims = (4, 6, 4, 8)
    d = gf.DeviceBuffer.from_numpy(_cfg(gf, dims))
    res = gf.run_mixed(gf.Lattice(dims), d, LANDAU, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=60,
                       mode=gf.MIXED_GAUGE24, check32=5)
    assert res.sweeps == res.sweeps_fp32 == 60
    assert res.checks == 2


def _oracle_coulomb_slices(oracle, dims, U4, precision=1e-12, maxiter=20000):
    """Each timeslice fixed by the oracle on its own, OR with omega 1.7,
    checked every sweep (the reference's per-slice loop,
    CoulombGaugeFixingSU3_4D.cu:198-317, with checkprecision = 1; the time
    links do not enter the Coulomb functional, so the slices are independent)."""
    dims3 = (1,) + tuple(dims[1:])
    S = to_timeslices(U4, dims)
    out = []
    for t in range(dims[0]):
        up = S[t].reshape(-1).copy()
        dw = S[t - 1].reshape(-1).copy()
        for it in range(maxiter):
            for parity in (0, 1):
                oracle.sweep_coulomb(OR, dims3, up, dw, parity, 1.7)
            g, th = oracle.quality(COULOMB, dims3, up)
            if th < precision:
                break
        out.append((g, th, it + 1))
    return out


def _first_sweep_below(oracle, gauge, dims, U, X, maxiter=1000):
    """The first FP64 OR sweep of the oracle (omega 1.7) after which theta <
    X; for Coulomb the largest timeslice theta, every slice swept on its own
    (the time links do not enter the Coulomb functional) in the order of the
    4D checkerboard: a site's parity there is that of t + x + y + z, so on an
    odd timeslice the even half-sweep updates the slice's odd sites first
    (the reference's per-slice loop sweeps every slice even-first; both
    orders reach the same maximum, along different paths)."""
    if gauge == COULOMB:
        dims3 = (1,) + tuple(dims[1:])
        S = to_timeslices(U, dims)
        slices = [(S[t].reshape(-1).copy(), S[t - 1].reshape(-1).copy()) for t in range(dims[0])]
    else:
        Uo = U.copy()
    for n in range(1, maxiter + 1):
        if gauge == COULOMB:
            th = 0.0
            for t, (up, dw) in enumerate(slices):
                for parity in (t % 2, 1 - t % 2):
                    oracle.sweep_coulomb(OR, dims3, up, dw, parity, 1.7)
                th = max(th, oracle.quality(COULOMB, dims3, up)[1])
        else:
            for parity in (0, 1):
                oracle.sweep(gauge, OR, dims, Uo, parity, 1.7)
            th = oracle.quality(gauge, dims, Uo.copy())[1]
        if th < X:
            return n
    raise AssertionError(f"the oracle's theta stays above {X} for {maxiter} sweeps")


@pytest.mark.parametrize("gauge,mode", [(LANDAU, 0), (LANDAU, 1), (LANDAU, 2), (LANDAU, 3), (COULOMB, 1), (COULOMB, 3)])
def test_fp32_check_ends_phase_where_fp64_theta_crosses(gf, gpu, oracle, gauge, mode):
    """The FP32 phase's fused check -- theta_x from the update's own matrix
    M = S g(x)^dagger in the gauge modes, from the links in the link mode --
    ends the phase at the first sweep whose theta is below switch_precision:
    with a check every sweep, within one sweep of where the FP64 oracle's
    theta (the reference's quality, computed after every sweep) first falls
    below it.  theta falls by 11-14% per sweep here (crossings at sweeps 93
    and 82), so a wrong term, factor or sign would move the crossing by
    several sweeps.  For Coulomb the check takes the largest timeslice theta;
    at 6 x 4 x 8 a timeslice's sites of one parity (96) are not a whole
    number of 64-site blocks, so each slice gets blocks of its own."""
    dims = (4, 6, 4, 8)
    U = _cfg(gf, dims)
    X = 1e-6
    n_star = _first_sweep_below(oracle, gauge, dims, U, X)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_mixed(gf.Lattice(dims), d, gauge, 1.7, switch_precision=X, precision=1e-30, maxiter=n_star + 20,
                       mode=mode, check32=1)
    print("crossing", gauge, mode, res.sweeps_fp32, n_star, flush=True)
    assert abs(res.sweeps_fp32 - n_star) <= 1, (res.sweeps_fp32, n_star)


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (4, 8, 8, 16), (16, 4, 4, 8)])
@pytest.mark.parametrize("driver", ["fast", "mixed", "mixed24"])
def test_coulomb_all_slices_at_once(gf, gpu, oracle, driver, dims):
    """Coulomb gauge on small 4D lattices, all timeslices in one run (each slice must meet theta < 1e-12), against the
    oracle fixing slice after slice: every slice's functional to 1e-10, every
    slice converged.  The shapes take the gauge-transformation sweep's
    timeslice-walking XCD block order with one and two timeslices per XCD
    (8^4, 16 x 4 x 4 x 8) and its plain XCD order (Nt = 4) over blocks laid out
    per timeslice."""
    V = int(np.prod(dims))
    U = _cfg(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    lat = gf.Lattice(dims)
    if driver == "fast":
        res = gf.run_fast(lat, d, gf.COULOMB, 1.7, 1e-12, 20000, 100)
    else:
        mode = gf.MIXED_GAUGE24 if driver == "mixed24" else gf.MIXED_GAUGE32
        res = gf.run_mixed(lat, d, gf.COULOMB, 1.7, switch_precision=1e-12, precision=1e-12, maxiter=20000,
                           mode=mode)
    assert res.converged == 1
    want = _oracle_coulomb_slices(oracle, dims, U)
    got = to_timeslices(d.download(), dims)
    dims3 = (1,) + dims[1:]
    for t in range(dims[0]):
        g, th = oracle.quality(COULOMB, dims3, got[t].reshape(-1).copy())
        assert g == pytest.approx(want[t][0], rel=1e-10), (t, g, want[t])
        assert th < 1e-12, (t, th)
    assert res.gff == pytest.approx(np.mean([w[0] for w in want]), rel=1e-10)


def _returned_theta(oracle, gauge, dims, Ud):
    """The oracle's theta of a returned lattice: the largest timeslice theta
    for Coulomb, the lattice's theta otherwise."""
    if gauge != COULOMB:
        return oracle.quality(gauge, dims, Ud.copy())[1]
    dims3 = (1,) + tuple(dims[1:])
    return max(oracle.quality(COULOMB, dims3, sl.reshape(-1).copy())[1] for sl in to_timeslices(Ud, dims))


@pytest.mark.parametrize("gauge,prec", [(LANDAU, 1e-12), (COULOMB, 1e-12), (MAG, 1e-10)])
@pytest.mark.parametrize("driver", ["fast", "mixed24", "mixed32", "links", "fp64gauge"])
def test_driver_returns_the_checked_state(gf, gpu, oracle, driver, gauge, prec):
    """The fused drivers hand back the configuration whose check met the
    precision, as the reference loop does (it breaks right after that check and
    projects, LandauGaugeFixingSU3_4D.cu:254-271): the oracle's theta of the
    returned lattice is below the precision and is the theta the driver
    reported (to the final projectSU3's rounding), and no reprojection enqueued
    after the converged sweep has touched it.  The FP64 phase forms its last
    checks with a read-only theta pass before the even update; a check fused
    into the even kernel would return the checked state plus one even
    half-sweep.  reproject = 7 puts a reprojection after most sweeps' ends."""
    if driver == "links" and gauge == COULOMB:
        pytest.skip("Coulomb runs with the gauge-transformation modes")
    dims = (4, 6, 4, 8)
    U = _cfg(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    lat = gf.Lattice(dims)
    maxiter = 20000
    if driver == "fast":
        res = gf.run_fast(lat, d, gauge, 1.7, prec, maxiter, 7)
    else:
        mode = {"mixed24": gf.MIXED_GAUGE24, "mixed32": gf.MIXED_GAUGE32, "links": gf.MIXED_LINKS,
                "fp64gauge": gf.MIXED_FP64}[driver]
        res = gf.run_mixed(lat, d, gauge, 1.7, switch_precision=1e-6, precision=prec, maxiter=maxiter, mode=mode,
                           check32=5, reproject=7)
    assert res.converged == 1, res
    th = _returned_theta(oracle, gauge, dims, d.download())
    print(driver, gauge, res, th, flush=True)
    assert th < prec, (th, res.theta)
    assert th == pytest.approx(res.theta, rel=1e-3), (th, res.theta)


@pytest.mark.parametrize("gauge,kw", [("landau", dict(sasteps=12, samax=0.4, samin=0.01)),
                                      ("u1xu1", dict(sasteps=10, samax=2.8, samin=1.9)),
                                      ("mag", dict(sasteps=4, samax=0.4, samin=0.01))])
def test_run_sa_schedule(gf, gpu, oracle, gauge, kw):
    """gf_run_sa against the apps' annealing schedule: the same temperatures
    and RNG counters as the oracle's schedule, and the identical result to the
    schedule driven step by step through gf_sweep / projectSU3 on the GPU.
    For Landau at these temperatures the following OR loop also ends on the
    oracle's functional to 1e-9."""
    from app_schedule import GAUGES, GpuBackend, Options, run_4d

    dims = (8, 8, 8, 8)
    U = _cfg(gf, dims)
    o = Options(randomtrafo=0, srmaxiter=0, ormaxiter=300, precision=1e-12, checkprecision=20, reproject=100, **kw)
    g = GAUGES[gauge]
    lat = gf.Lattice(dims)
    # the schedule alone (no OR) driven through the library
    o_sa = Options(**{**o.__dict__, "ormaxiter": 0})
    ref_gpu = run_4d(GpuBackend(gf), gauge, dims, U, o_sa)[0]
    d = gf.DeviceBuffer.from_numpy(U)
    ctr, ntemps = gf.run_sa(lat, d, g, o.samax, o.samin, o.sasteps, o.microupdates, o.reproject, o.seed, 0)
    assert ctr == ref_gpu["counter"] and ntemps > 0
    gf.CommonKernelsSU3(lat).projectSU3(lat.V // 32, 32, d)  # the schedule's final projectSU3
    assert np.array_equal(d.download(), ref_gpu["U"])
    want = run_4d(oracle, gauge, dims, U, o)[0]
    assert ctr == want["counter"]
    if g == LANDAU:
        d = gf.DeviceBuffer.from_numpy(U)
        gf.run_sa(lat, d, g, o.samax, o.samin, o.sasteps, o.microupdates, o.reproject, o.seed, 0)
        res = gf.run_or(lat, d, g, o.orparameter, o.precision, o.ormaxiter, o.checkprecision, o.reproject)
        assert res.sweeps == want["or_sweeps"]
        assert res.gff == pytest.approx(want["gff"], rel=1e-9)


@pytest.mark.parametrize("mode", [1, 2])
def test_run_sa_mag_gauge_modes(gf, gpu, oracle, mode):
    """MAG's annealing schedule on the gauge transformation (five heatbath
    steps per temperature): the same number of temperatures and RNG counters
    as the link sweeps, the links change by a gauge transformation only
    (plaquette to 1e-13), and the functional rises as with the link sweeps."""
    from oracle_lib import to_matrices
    from test_gpu_parity import _plaquette

    dims = (8, 8, 8, 8)
    V = int(np.prod(dims))
    U = _cfg(gf, dims, sigma=0.35)
    lat = gf.Lattice(dims)
    g0, _ = oracle.quality(MAG, dims, U.copy())
    out = {}
    for m in (0, mode):
        d = gf.DeviceBuffer.from_numpy(U)
        out[m] = gf.run_sa(lat, d, gf.MAG, 0.9, 0.6, 20, 2, 100, 1, 0, mode=m)
        gf.CommonKernelsSU3(lat).projectSU3(lat.V // 32, 32, d)
        Ua = d.download()
        if m != 0:
            assert _plaquette(to_matrices(Ua, V), dims) == pytest.approx(_plaquette(to_matrices(U, V), dims), rel=1e-13)
        out[m] = out[m] + (oracle.quality(MAG, dims, Ua.copy())[0],)
    assert out[mode][:2] == out[0][:2]
    assert out[mode][2] > g0 + 0.02 and out[0][2] > g0 + 0.02
    assert out[mode][2] == pytest.approx(out[0][2], abs=0.02)


def test_run_sa_gauge_mode_is_a_gauge_transformation(gf, gpu, oracle):
    """Annealing the gauge transformation in FP32 (gf_run_sa mode 1, and mode 2
    on the links in 24-bit fixed point) changes the links only by a gauge
    transformation applied in FP64: the plaquette
    is unchanged to 1e-13 and the links stay unitary; the functional rises
    like the reference-precision annealing's, and the OR loop then converges."""
    from test_gpu_parity import _plaquette
    from oracle_lib import to_matrices

    dims = (8, 8, 8, 8)
    V = int(np.prod(dims))
    U = _cfg(gf, dims)
    lat = gf.Lattice(dims)
    g0, _ = oracle.quality(LANDAU, dims, U.copy())
    res = {}
    for mode in (0, 1, 2):
        d = gf.DeviceBuffer.from_numpy(U)
        ctr, ntemps = gf.run_sa(lat, d, gf.LANDAU, 1.45, 0.01, 40, 3, 100, 1, 0, mode=mode)
        assert (ctr, ntemps) == (80, 40)
        gf.CommonKernelsSU3(lat).projectSU3(lat.V // 32, 32, d)  # third rows (mode 0 leaves them stale)
        Ua = d.download()
        M = to_matrices(Ua, V)
        assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-13
        res[mode] = oracle.quality(LANDAU, dims, Ua.copy())[0]
        if mode != 0:
            assert _plaquette(M, dims) == pytest.approx(_plaquette(to_matrices(U, V), dims), rel=1e-13)
            r = gf.run_or(lat, d, gf.LANDAU, 1.7, 1e-12, 5000, 20, 100)
            assert r.converged == 1
    assert res[0] > g0 + 0.05 and res[1] > g0 + 0.05 and res[2] > g0 + 0.05
    assert res[1] == pytest.approx(res[0], abs=0.02)
    assert res[2] == pytest.approx(res[0], abs=0.02)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_run_mixed_mag(gf, gpu, oracle, mode):
    """MAG through the mixed driver, in the link mode (FP32 links with g) and
    on the gauge transformation (FP32 / 24-bit links, FP32 / FP64 arithmetic)
    until theta < 1e-6 or stagnation, then FP64 to 1e-10: it ends on the FP64
    oracle's maximum, functional to 1e-10, and the result is a gauge
    transformation of the input (plaquette to 1e-13)."""
    from oracle_lib import to_matrices
    from test_gpu_parity import _plaquette

    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    U = _cfg(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    res = gf.run_mixed(gf.Lattice(dims), d, gf.MAG, 1.7, switch_precision=1e-6, precision=1e-10, maxiter=20000,
                       mode=mode, check32=10)
    Uo = U.copy()
    n, g, t = oracle.run_or(MAG, dims, Uo, 1.7, 1e-10, 20000, 1, 100)
    assert res.converged == 1 and res.theta < 1e-10
    assert 0 < res.sweeps_fp32 < res.sweeps
    assert res.gff == pytest.approx(g, rel=1e-10)
    assert _plaquette(to_matrices(d.download(), V), dims) == pytest.approx(_plaquette(to_matrices(U, V), dims),
                                                                           rel=1e-13)


def test_mag_fp64_phase_coarse_checks_keep_the_sweep_count(gf, gpu):
    """MAG's FP64 phase checks every 16 sweeps while theta, falling at the
    rate of the host's reads, is more than 160 sweeps from the target.  On a
    12^4 lattice (sigma 0.35) taken to theta < 1e-12 -- thousands of sweeps,
    most of them with coarse checks -- the all-FP64 gauge-transformation run
    still stops on the sweep count of run_fast, which checks every sweep
    (+-1: the two round differently), and the returned links, measured again
    by the drop-in quality kernel, meet the precision; so does the mixed run
    (24-bit FP32 phase to 1e-6), on the same functional."""
    dims = (12, 12, 12, 12)
    lat = gf.Lattice(dims)
    U = _cfg(gf, dims, sigma=0.35)

    def theta_of(buf):
        st = gf.GaugeFixingStats(gf.MAGKernelsSU3, buf, lat)
        st.generateGaugeQuality()
        t = st.getCurrentA()
        st.close()
        return t

    d_fast = gf.DeviceBuffer.from_numpy(U)
    fast = gf.run_fast(lat, d_fast, gf.MAG, 1.7, 1e-12, 200000, 100)
    d_g64 = gf.DeviceBuffer.from_numpy(U)
    g64 = gf.run_mixed(lat, d_g64, gf.MAG, 1.7, 1e-12, 1e-12, 200000, mode=gf.MIXED_FP64)
    d_mix = gf.DeviceBuffer.from_numpy(U)
    mix = gf.run_mixed(lat, d_mix, gf.MAG, 1.7, 1e-6, 1e-12, 200000, mode=gf.MIXED_GAUGE24)
    assert fast.converged == g64.converged == mix.converged == 1
    assert fast.sweeps > 1000
    assert abs(g64.sweeps - fast.sweeps) <= 1, (g64.sweeps, fast.sweeps)
    assert g64.checks < fast.sweeps / 4, (g64.checks, fast.sweeps)  # coarse checks took most of the run
    for buf, r in ((d_fast, fast), (d_g64, g64), (d_mix, mix)):
        assert r.theta < 1e-12
        assert theta_of(buf) < 1e-12
    assert g64.gff == pytest.approx(fast.gff, rel=1e-12)
    assert mix.gff == pytest.approx(fast.gff, rel=1e-12)


@pytest.mark.parametrize("gauge,prec", [(LANDAU, 1e-12), (COULOMB, 1e-12), (MAG, 1e-10), (U1XU1, 1e-30)])
def test_fp64_gauge_sweep_matches_fp64_link_sweeps(gf, gpu, oracle, gauge, prec):
    """The all-FP64 run on the FP64 gauge-transformation sweep (run_mixed
    without its FP32 phase: a gauge transformation g64 swept on the links held
    read only, g64 U g64^dagger written back) makes the OR updates of the FP64
    link sweep, rounded differently: the same sweep count to +-1 as run_fast
    and the oracle's run_or, the oracle's functional to 1e-12, no FP32 sweep,
    and the result a gauge transformation of the input (plaquette to 1e-13).
    U1xU1 runs to a fixed sweep count (its theta need not converge) and is
    compared with run_fast's links after the same sweeps."""
    from oracle_lib import to_matrices
    from test_gpu_parity import _plaquette

    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    U = _cfg(gf, dims)
    lat = gf.Lattice(dims)
    d = gf.DeviceBuffer.from_numpy(U)
    maxiter = 400 if gauge == U1XU1 else 20000
    res = gf.run_mixed(lat, d, gauge, 1.7, switch_precision=1e-6, precision=prec, maxiter=maxiter,
                       mode=gf.MIXED_FP64, reproject=100)
    d2 = gf.DeviceBuffer.from_numpy(U)
    f = gf.run_fast(lat, d2, gauge, 1.7, prec, maxiter, 100)
    print(gauge, res, f, flush=True)
    assert res.sweeps_fp32 == 0
    assert res.converged == f.converged
    if gauge == U1XU1:
        assert res.sweeps == f.sweeps == maxiter
        assert np.abs(rows01(d.download(), V) - rows01(d2.download(), V)).max() < 1e-10
    else:
        assert res.converged == 1 and res.theta < prec
        assert abs(res.sweeps - f.sweeps) <= 1
        if gauge != COULOMB:
            Uo = U.copy()
            n, g, t = oracle.run_or(gauge, dims, Uo, 1.7, prec, maxiter, 1, 100)
            assert abs(res.sweeps - n) <= 1
            assert res.gff == pytest.approx(g, rel=1e-12)
    assert res.gff == pytest.approx(f.gff, rel=1e-12)
    assert _plaquette(to_matrices(d.download(), V), dims) == pytest.approx(_plaquette(to_matrices(U, V), dims),
                                                                           rel=1e-13)


@pytest.mark.parametrize("driver", ["fast", "mixed24"])
def test_unconverged_coulomb_reports_largest_slice_theta(gf, gpu, oracle, driver):
    """A Coulomb run on a 4D lattice that stops at maxiter reports the theta
    its checks test -- the largest timeslice theta -- for the returned
    lattice, not the average over the slices, and the slices' average
    functional."""
    dims = (4, 6, 4, 8)
    U = _cfg(gf, dims)
    d = gf.DeviceBuffer.from_numpy(U)
    lat = gf.Lattice(dims)
    if driver == "fast":
        res = gf.run_fast(lat, d, COULOMB, 1.7, 1e-12, 20, 100)
    else:
        res = gf.run_mixed(lat, d, COULOMB, 1.7, switch_precision=1e-6, precision=1e-12, maxiter=20, check32=5)
    assert res.converged == 0 and res.sweeps == 20
    dims3 = (1,) + dims[1:]
    q = [oracle.quality(COULOMB, dims3, sl.reshape(-1).copy()) for sl in to_timeslices(d.download(), dims)]
    assert res.theta == pytest.approx(max(t for _, t in q), rel=1e-9)
    assert abs(res.theta - np.mean([t for _, t in q])) > 1e-3 * res.theta
    assert res.gff == pytest.approx(np.mean([g for g, _ in q]), rel=1e-12)


def test_u1xu1_on_mag_result(gf, gpu, oracle):
    """The MAG workload's chain on a small lattice (sigma 0.35): MAG fixed in
    mixed precision to theta < 1e-10 (the oracle's maximum, functional to
    1e-10), then 100 U1xU1 OR sweeps on the result by the fused FP64 driver
    and by the oracle from the same start (both reproject at sweep 0): the
    U1xU1 functionals agree to 1e-10, and neither run changes MAG's
    functional or theta, since a diagonal gauge rotation changes neither
    |U_jj| nor |X_k|."""
    dims = (8, 8, 8, 8)
    V = int(np.prod(dims))
    U = _cfg(gf, dims, sigma=0.35)
    lat = gf.Lattice(dims)
    d = gf.DeviceBuffer.from_numpy(U)
    m = gf.run_mixed(lat, d, gf.MAG, 1.7, switch_precision=1e-6, precision=1e-10, maxiter=50000, check32=5)
    n, g_or, _ = oracle.run_or(MAG, dims, U.copy(), 1.7, 1e-10, 50000, 1, 100)
    assert m.converged == 1 and m.gff == pytest.approx(g_or, rel=1e-10)
    R = d.download()
    g_mag, t_mag = oracle.quality(MAG, dims, R.copy())
    assert t_mag < 1e-10
    u = gf.run_fast(lat, d, gf.U1XU1, 1.7, 0.0, 100, 100)
    Ro = R.copy()
    no, g_uo, _ = oracle.run_or(U1XU1, dims, Ro, 1.7, 0.0, 100, 1, 100)
    assert u.sweeps == no == 100 and not u.converged
    assert u.gff == pytest.approx(g_uo, rel=1e-10)
    g_u1_start = oracle.quality(U1XU1, dims, R.copy())[0]
    assert u.gff > g_u1_start + 1e-4  # the U1xU1 sweeps did work
    for after in (d.download(), Ro):
        g2, t2 = oracle.quality(MAG, dims, after.copy())
        assert g2 == pytest.approx(g_mag, rel=1e-12)
        assert t2 == pytest.approx(t_mag, rel=1e-3, abs=1e-13)
    del V


@pytest.mark.parametrize("gauge,mode", [(LANDAU, 0), (LANDAU, 1), (LANDAU, 2), (MAG, 2), (U1XU1, 0)])
def test_run_sa_graph_is_bit_identical(gf, gpu, gauge, mode):
    """GF_SA_GRAPH: the annealing schedule captured into one HIP graph and
    replayed gives the links, counters and temperature count of the same
    schedule launched one kernel at a time -- on the default stream (the call
    captures on a stream of its own) and on the caller's stream, and through
    the batch call on the caller's streams."""
    import ctypes as C

    dims = (8, 8, 8, 8)
    U = _cfg(gf, dims, sigma=0.35 if gauge == MAG else 0.3)
    lat = gf.Lattice(dims)
    kw = dict(tmax=0.9, tmin=0.6, steps=20) if gauge == MAG else dict(tmax=1.45, tmin=0.01, steps=40)
    if gauge == U1XU1:
        kw = dict(tmax=2.8, tmin=1.9, steps=10)
    stream = C.c_void_p()
    gf._lib.call("gf_stream_create", C.byref(stream))
    out = {}
    for graph, st in ((False, None), (True, None), (True, stream)):
        d = gf.DeviceBuffer.from_numpy(U)
        ctr = gf.run_sa(lat, d, gauge, micro=3, reproject=7, seed=5, counter=100, mode=mode, stream=st, graph=graph,
                        **kw)
        gf.synchronize()
        out[(graph, st is not None)] = (ctr, d.download())
    ref = out[(False, False)]
    for k, (ctr, u) in out.items():
        assert ctr == ref[0], k
        assert np.array_equal(u, ref[1]), k
    stream2 = C.c_void_p()
    gf._lib.call("gf_stream_create", C.byref(stream2))
    bufs = [gf.DeviceBuffer.from_numpy(U) for _ in range(2)]
    res = gf.batch_run_sa([lat] * 2, bufs, gauge, micro=3, reproject=7, seeds=[5, 5], counters=[100, 100], mode=mode,
                          max_concurrent=2, streams=[stream, stream2], graph=True, **kw)
    gf.synchronize()
    for r, b in zip(res, bufs):
        assert r == ref[0]
        assert np.array_equal(b.download(), ref[1])
    for st in (stream, stream2):
        gf._lib.call("gf_stream_destroy", st)


def _set_wavefront(gf, slab):
    import ctypes as C
    prev = C.c_int()
    gf._lib.call("gf_set_wavefront", int(slab), C.byref(prev))
    return prev.value


@pytest.mark.parametrize("gauge,dims", [(LANDAU, (8, 8, 8, 8)), (LANDAU, (12, 8, 8, 8)), (COULOMB, (8, 6, 4, 8)),
                                        (MAG, (8, 8, 8, 8)), (U1XU1, (6, 8, 8, 8))])
def test_wavefront_sweeps_are_bit_identical(gf, gpu, gauge, dims):
    """A sweep run as a wavefront over timeslices (launch k: even sites of
    slab k, odd sites of the slab before it shifted back by one timeslice,
    timeslice 0's odd sites last) gives the two-kernel sweep's result bit for
    bit, for every slab from one timeslice to more than the lattice: the mixed
    driver's FP32 phase (24-bit and FP32 links; sweeps without a check) and
    the gauge-annealing schedule (heatbath random numbers per site and
    parity, microcanonical steps)."""
    U = _cfg(gf, dims, sigma=0.35 if gauge == MAG else 0.3)
    lat = gf.Lattice(dims)
    prev = _set_wavefront(gf, 0)
    try:
        want = {}
        for slab in (0, 1, 2, 3, 5, dims[0] + 1):
            _set_wavefront(gf, slab)
            got = {}
            for mode in (gf.MIXED_GAUGE24, gf.MIXED_GAUGE32):
                d = gf.DeviceBuffer.from_numpy(U)
                r = gf.run_mixed(lat, d, gauge, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=23, mode=mode,
                                 check32=5)
                got[mode] = (r.sweeps, r.checks, d.download())
            if gauge != COULOMB:
                d = gf.DeviceBuffer.from_numpy(U)
                ctr = gf.run_sa(lat, d, gauge, 1.2, 0.1, 6, 2, 100, 3, 40, mode=2)
                got["sa"] = (ctr, d.download())
            if slab == 0:
                want = got
                continue
            for k, v in got.items():
                assert v[:-1] == want[k][:-1], (slab, k)
                assert np.array_equal(v[-1], want[k][-1]), (slab, k)
    finally:
        _set_wavefront(gf, prev)


def _set_block_order(gf, tpass):
    import ctypes as C
    prev = C.c_int()
    gf._lib.call("gf_set_block_order", int(tpass), C.byref(prev))
    return prev.value


@pytest.mark.parametrize("gauge,dims", [(LANDAU, (16, 8, 8, 8)), (LANDAU, (24, 4, 8, 8)), (COULOMB, (16, 4, 8, 8)),
                                        (MAG, (16, 8, 8, 8))])
def test_block_order_is_bit_identical(gf, gpu, gauge, dims):
    """The XCD block order of the gauge-transformation sweep (each XCD's
    timeslices walked in passes of 1, 2, 3 or all of them) changes which
    block runs where and when, never a result: the mixed driver's run (FP32
    phase with checks every 5 sweeps on 24-bit and FP32 links, FP32 phase in
    FP64 arithmetic, and the FP64 phase on the FP64 gauge-transformation
    sweep, all-FP64 mode) and the gauge-annealing schedule are bit-identical
    for every order.  The shapes give each XCD 2 or 3 timeslices of whole
    64-site blocks."""
    U = _cfg(gf, dims, sigma=0.35 if gauge == MAG else 0.3)
    lat = gf.Lattice(dims)
    spx = dims[0] // 8
    prev_wf = _set_wavefront(gf, 0)
    prev = _set_block_order(gf, 0)
    try:
        want = None
        for tpass in [0] + [p for p in range(1, spx) if spx % p == 0]:
            _set_block_order(gf, tpass)
            got = {}
            for mode in (gf.MIXED_GAUGE24, gf.MIXED_GAUGE32, gf.MIXED_GAUGE64, gf.MIXED_FP64):
                d = gf.DeviceBuffer.from_numpy(U)
                r = gf.run_mixed(lat, d, gauge, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=23, mode=mode,
                                 check32=5)
                got[mode] = ((r.sweeps, r.checks, r.theta, r.gff), d.download())
            if gauge != COULOMB:
                d = gf.DeviceBuffer.from_numpy(U)
                ctr = gf.run_sa(lat, d, gauge, 1.2, 0.1, 6, 2, 100, 3, 40, mode=2)
                got["sa"] = ((ctr,), d.download())
            if want is None:
                want = got
                continue
            for k, v in got.items():
                assert v[0] == want[k][0], (tpass, k)
                assert np.array_equal(v[1], want[k][1]), (tpass, k)
    finally:
        _set_block_order(gf, prev)
        _set_wavefront(gf, prev_wf)


def _pipelined_launches(gf):
    import ctypes as C
    n = C.c_longlong()
    gf._lib.call("gf_pipelined_launches", C.byref(n))
    return n.value


def _set_pipelined_lag(gf, lag):
    import ctypes as C
    prev = C.c_int()
    gf._lib.call("gf_set_pipelined_lag", int(lag), C.byref(prev))
    return prev.value


@pytest.mark.parametrize("gauge,dims", [(LANDAU, (8, 8, 8, 8)), (LANDAU, (16, 8, 12, 16)), (LANDAU, (24, 8, 8, 8)),
                                        (U1XU1, (16, 16, 16, 16)), (LANDAU, (32, 32, 32, 32)),
                                        (COULOMB, (8, 8, 8, 16)), (COULOMB, (16, 8, 12, 16)),
                                        (COULOMB, (64, 32, 32, 32))])
def test_pipelined_sweeps_are_bit_identical(gf, gpu, gauge, dims):
    """A sweep run as the pipelined sweep (one launch: each XCD's queue of
    even and odd 64-site blocks, an odd block starting once the even blocks
    around it have set their flags) gives the two-kernel sweep's result bit
    for bit: the mixed driver's FP32 phase on 24-bit and FP32 links in FP32
    and FP64 arithmetic (the sweeps between checks every 5 sweeps), and the
    gauge-annealing schedule on 24-bit and FP32 links (heatbath random
    numbers per site and parity, microcanonical steps), with no extra lag and
    with some.  The shapes give each queue one to eight timeslices, include a
    neighbour reach that is not a whole number of blocks (16x8x12x16: 96
    sites), the benchmark's lattice and Coulomb's 32^3 x 64 (every timeslice
    at once, its odd blocks waiting on their own timeslice only); the
    library's launch count shows that the pipelined sweep ran."""
    U = _cfg(gf, dims)
    lat = gf.Lattice(dims)
    # every setting fixed, so that no timing of candidates launches sweeps of its own
    prev = _set_wavefront(gf, 0)
    prev_lag = _set_pipelined_lag(gf, 0)
    prev_order = _set_block_order(gf, 0)
    try:
        want = None
        for slab, lag in ((0, 0), (gf.WAVEFRONT_PIPELINED, 0), (gf.WAVEFRONT_PIPELINED, 37),
                          (gf.WAVEFRONT_PIPELINED, 384)):
            _set_wavefront(gf, slab)
            _set_pipelined_lag(gf, lag)
            n0 = _pipelined_launches(gf)
            got = {}
            for mode in (gf.MIXED_GAUGE24, gf.MIXED_GAUGE32, gf.MIXED_GAUGE64):
                d = gf.DeviceBuffer.from_numpy(U)
                r = gf.run_mixed(lat, d, gauge, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=23, mode=mode,
                                 check32=5)
                got[mode] = ((r.sweeps, r.checks, r.theta, r.gff), d.download())
            for sa_mode in (1, 2):
                d = gf.DeviceBuffer.from_numpy(U)
                ctr = gf.run_sa(lat, d, gauge, 1.2, 0.1, 6, 2, 100, 3, 40, mode=sa_mode)
                got["sa", sa_mode] = ((ctr,), d.download())
            launched = _pipelined_launches(gf) - n0
            print(dims, slab, lag, "pipelined launches", launched, flush=True)
            if slab == 0:
                assert launched == 0
                want = got
                continue
            # 3 mixed runs x 23 sweeps (checks every 5) + 2 schedules x 6 x (1 + 2) sweeps
            assert launched >= 3 * 18 + 2 * 18
            for k, v in got.items():
                assert v[0] == want[k][0], (slab, lag, k)
                assert np.array_equal(v[1], want[k][1]), (slab, lag, k)
    finally:
        _set_wavefront(gf, prev)
        _set_pipelined_lag(gf, prev_lag)
        _set_block_order(gf, prev_order)
