"""CPU checks of the pieces the application test relies on: the timeslice
re-layout, the option parser round trip and the annealing schedule's float
arithmetic (which decides how many heatbath steps run)."""
import numpy as np

from app_schedule import F, Options, parse_results, to_timeslices


def test_to_timeslices_matches_site_indices(oracle):
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    dims3 = (1,) + dims[1:]
    V3 = V // dims[0]
    U = np.arange(72 * V, dtype=np.float64)
    S = to_timeslices(U, dims)
    for t in range(dims[0]):
        for n3 in (0, 1, 7, V3 // 2, V3 - 1):
            _, x, y, z = oracle.site_coords(dims3, n3)
            i4 = oracle.site_index(dims, (t, x, y, z))
            assert np.array_equal(S[t, :, n3], U.reshape(72, V)[:, i4])


def test_options_argv_round_trip():
    o = Options(samax=2.8, samin=1.9, precision=1e-12)
    argv = o.argv()
    kv = dict(zip(argv[::2], argv[1::2]))
    assert float(np.float32(float(kv["--samax"]))) == float(F(2.8))
    assert kv["--precision"] == "1e-12" and kv["--gaugecopies"] == "1"


def test_parse_results():
    out = "x\nRESULT copy=0 gff=0.5 theta=1e-13 or_sweeps=12 sr_sweeps=0 counter=4\nBEST gff=0.5\n"
    assert parse_results(out) == [dict(copy=0, gff=0.5, theta=1e-13, or_sweeps=12, sr_sweeps=0, counter=4)]


def test_u1xu1_schedule_temperatures():
    """do { step = fine if 2 <= T <= 2.7 else coarse; ...; T -= step } while T >= Tmin,
    with float T and double bounds (U1xU1GaugeFixingSU3_4D.cu:219-251)."""
    base = F(F(2.8) - F(1.9)) / F(10)
    T, temps = F(2.8), []
    while True:
        temps.append(float(T))
        step = F(base / F(15)) if 2.0 <= float(T) <= 2.7 else base
        T = F(T - step)
        if not T >= F(1.9):
            break
    fine = [t for t in temps if 2.0 <= t <= 2.7]
    assert len(temps) > 100 and len(fine) > 95
    assert all(a > b for a, b in zip(temps, temps[1:]))
