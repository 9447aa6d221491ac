"""The benchmark's own lattices through size-independent properties: 32^4
Landau (sigma 0.25, seed 1, Haar seed 2), the 32^3 x 64 Coulomb lattice and
the gauge copies' 16^4 to 32^4 lattices with the full annealing schedule.
The CPU oracle would need minutes for one run at these sizes, so the checks
compare the GPU paths with each other (the oracle only measures the results).

* The fused FP64 driver (check every sweep) stops within one sweep of the
  drop-in reference loop (gf_run_or: the reference-order kernels, which are
  bit-identical to the reference's own code, and its quality check after
  every sweep), on the same functional to 1e-12.
* The mixed-precision run ends on the FP64 run's functional to 1e-10, with
  theta below 1e-12, and its result is a gauge transformation of the input:
  the average plaquette, which no gauge transformation changes, is the
  input's to 1e-12.
"""
import numpy as np
import pytest

from oracle_lib import to_matrices
from test_gpu_parity import _plaquette

pytestmark = pytest.mark.gpu

DIMS = (32, 32, 32, 32)


@pytest.fixture(scope="module")
def config(gf):
    print("generating the 32^4 configuration", flush=True)
    return gf.make_config(DIMS, 0.25, 1, 2)


def test_fused_driver_stops_with_reference_loop(gf, gpu, config):
    lat = gf.Lattice(DIMS)
    d = gf.DeviceBuffer.from_numpy(config)
    ref = gf.run_or(lat, d, gf.LANDAU, 1.7, 1e-12, 5000, 1, 100)
    print("run_or", ref, flush=True)
    d = gf.DeviceBuffer.from_numpy(config)
    fast = gf.run_fast(lat, d, gf.LANDAU, 1.7, 1e-12, 5000, 100)
    print("run_fast", fast, flush=True)
    assert ref.converged and fast.converged
    assert abs(fast.sweeps - ref.sweeps) <= 1, (fast.sweeps, ref.sweeps)
    assert fast.gff == pytest.approx(ref.gff, rel=1e-12)


def test_mixed_run_on_fp64_maximum_and_gauge_orbit(gf, gpu, config):
    V = int(np.prod(DIMS))
    lat = gf.Lattice(DIMS)
    d = gf.DeviceBuffer.from_numpy(config)
    fp64 = gf.run_fast(lat, d, gf.LANDAU, 1.7, 1e-12, 5000, 100)
    d = gf.DeviceBuffer.from_numpy(config)
    mixed = gf.run_mixed(lat, d, gf.LANDAU, 1.7, switch_precision=1e-12, precision=1e-12, maxiter=5000)
    print("fp64", fp64, "mixed", mixed, flush=True)
    assert mixed.converged and mixed.theta < 1e-12
    assert mixed.gff == pytest.approx(fp64.gff, rel=1e-10)
    p_in = _plaquette(to_matrices(config, V), DIMS)
    p_out = _plaquette(to_matrices(d.download(np.float64, config.size), V), DIMS)
    assert p_out == pytest.approx(p_in, rel=1e-12)


COULOMB_DIMS = (64, 32, 32, 32)


def test_coulomb_32x32x32x64_every_slice(gf, gpu, oracle):
    """The Coulomb workload's lattice (32^3 x 64, sigma 0.3, seed 1, Haar
    seed 2), every timeslice at once, in FP64 and in mixed precision -- the
    bench's schedule (24-bit links, switch at 1e-12) and the workload's stated
    one (FP32 12-real links until theta < 1e-6, then FP64): every slice of
    each result has theta < 1e-12 by the oracle's per-slice quality, the
    results agree slice by slice on the functional to 1e-10, and the mixed
    results are gauge transformations of the input (plaquette to 1e-12)."""
    from app_schedule import to_timeslices
    from oracle_lib import COULOMB

    V = int(np.prod(COULOMB_DIMS))
    dims3 = (1,) + COULOMB_DIMS[1:]
    print("generating the 32^3 x 64 configuration", flush=True)
    U = gf.make_config(COULOMB_DIMS, 0.3, 1, 2)
    lat = gf.Lattice(COULOMB_DIMS)
    slices = {}
    for name in ("fp64", "mixed", "mixed_fp32_1e-6"):
        d = gf.DeviceBuffer.from_numpy(U)
        if name == "fp64":
            r = gf.run_fast(lat, d, gf.COULOMB, 1.7, 1e-12, 20000, 100)
        elif name == "mixed":
            r = gf.run_mixed(lat, d, gf.COULOMB, 1.7, switch_precision=1e-12, precision=1e-12, maxiter=20000)
        else:
            r = gf.run_mixed(lat, d, gf.COULOMB, 1.7, switch_precision=1e-6, precision=1e-12, maxiter=20000,
                             mode=gf.MIXED_GAUGE32)
            assert 0 < r.sweeps_fp32 < r.sweeps
        print(name, r, flush=True)
        assert r.converged
        out = d.download(np.float64, U.size)
        S = to_timeslices(out, COULOMB_DIMS)
        slices[name] = [oracle.quality(COULOMB, dims3, S[t].reshape(-1).copy()) for t in range(COULOMB_DIMS[0])]
        assert max(th for _, th in slices[name]) < 1e-12, name
        if name != "fp64":
            p_in = _plaquette(to_matrices(U, V), COULOMB_DIMS)
            p_out = _plaquette(to_matrices(out, V), COULOMB_DIMS)
            assert p_out == pytest.approx(p_in, rel=1e-12)
    for name in ("mixed", "mixed_fp32_1e-6"):
        for (g64, _), (gm, _) in zip(slices["fp64"], slices[name]):
            assert gm == pytest.approx(g64, rel=1e-10), name


@pytest.mark.parametrize("L,copies", [(16, 4), (20, 4), (24, 2), (32, 1)])
def test_gauge_copies_full_schedule(gf, gpu, L, copies):
    """The gauge-copies workload's L^4 copies (sigma 0.3, seed 1, Haar seeds
    10.. one per copy, seed 1 with counter ranges 100000 apart) through the
    batch calls with the full schedule: annealing T 1.45 -> 0.01 in 3000 steps
    with 3 microcanonical steps each (FP32 gauge annealing on 24-bit links),
    then mixed OR to theta < 1e-12.  Every copy converges; the copies end on
    one functional to 1e-12 (this smooth configuration has one maximum); each
    result is a gauge transformation of its input (plaquette to 1e-12).  The
    workload's four copies of 16^4 and 20^4, two of 24^4 and one of 32^4 (the
    sizes at which a copy fills the GPU alone run one after another)."""
    dims = (L, L, L, L)
    V = L ** 4
    cfgs = [gf.make_config(dims, 0.3, 1, 10 + c) for c in range(copies)]
    bufs = [gf.DeviceBuffer.from_numpy(u) for u in cfgs]
    lats = [gf.Lattice(dims)] * copies
    sa = gf.batch_run_sa(lats, bufs, gf.LANDAU, 1.45, 0.01, 3000, 3, 100, seeds=[1] * copies,
                         counters=[c * 100000 for c in range(copies)], mode=2)
    assert all(t == 3000 for _, t in sa)
    res = gf.batch_run_mixed(lats, bufs, gf.LANDAU, 1.7, 1e-12, 1e-12, 20000, check32=5)
    print(L, [r.gff for r in res], flush=True)
    assert all(r.converged == 1 and r.theta < 1e-12 for r in res)
    gffs = [r.gff for r in res]
    assert max(gffs) - min(gffs) < 1e-12 * max(gffs)
    for u, b in zip(cfgs, bufs):
        p_in = _plaquette(to_matrices(u, V), dims)
        p_out = _plaquette(to_matrices(b.download(np.float64, u.size), V), dims)
        assert p_out == pytest.approx(p_in, rel=1e-12)


def test_mag_48_bounded(gf, gpu):
    """The MAG workload's 48^4 lattice (sigma 0.35, seed 1, Haar seed 2) over
    a bounded number of sweeps (the full run needs some 10^5): 200 mixed
    sweeps (all in the FP32 phase, on 24-bit links) and 200 FP64 sweeps raise
    MAG's functional by the same amount to FP32 accuracy, the mixed result is
    a gauge transformation of the input (the plaquette of every 97th site to
    1e-12), and 50 U1xU1 sweeps on the FP64 result keep MAG's functional and
    theta (diagonal gauge rotations)."""
    dims = (48, 48, 48, 48)
    V = 48 ** 4
    print("generating the 48^4 configuration", flush=True)
    U = gf.make_config(dims, 0.35, 1, 2)
    lat = gf.Lattice(dims)
    d = gf.DeviceBuffer.from_numpy(U)

    def mag_q(buf):
        st = gf.GaugeFixingStats(gf.MAGKernelsSU3, buf, lat)
        st.generateGaugeQuality()
        out = (st.getCurrentGff(), st.getCurrentA())
        st.close()
        return out

    g0, t0 = mag_q(d)
    m = gf.run_mixed(lat, d, gf.MAG, 1.7, switch_precision=1e-30, precision=1e-30, maxiter=200, check32=5)
    print("mixed", m, flush=True)
    assert m.sweeps == m.sweeps_fp32 == 200 and not m.converged
    d64 = gf.DeviceBuffer.from_numpy(U)
    f = gf.run_fast(lat, d64, gf.MAG, 1.7, 0.0, 200, 100)
    print("fp64", f, flush=True)
    assert f.sweeps == 200 and not f.converged
    gm, tm = mag_q(d)
    gf64, tf64 = mag_q(d64)
    assert gm > g0 + 0.1 and gf64 > g0 + 0.1
    assert gm == pytest.approx(gf64, rel=1e-6)
    assert tm == pytest.approx(tf64, rel=1e-2)
    p_in = _plaquette(to_matrices(U, V), dims, every=97)
    p_out = _plaquette(to_matrices(d.download(np.float64, U.size), V), dims, every=97)
    assert p_out == pytest.approx(p_in, rel=1e-12)
    u = gf.run_fast(lat, d64, gf.U1XU1, 1.7, 0.0, 50, 100)
    g1, t1 = mag_q(d64)
    print("u1xu1", u, g1, t1, flush=True)
    assert g1 == pytest.approx(gf64, rel=1e-12)
    assert t1 == pytest.approx(tf64, rel=1e-6)
