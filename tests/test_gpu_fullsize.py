"""The benchmark's own lattice, 32^4 Landau (sigma 0.25, seed 1, Haar seed 2),
through size-independent properties: the CPU oracle would need minutes for
one run at this size, so the checks compare the GPU paths with each other.

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
