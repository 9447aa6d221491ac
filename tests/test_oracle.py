"""Pins of the CPU oracle (the checker of the gfx950 kernels).

1. Golden vectors (tests/golden/ref_vectors.npz) written from the
   reference's own lattice and Random123 headers compiled by oracle/Makefile:
   geometry, Philox / u01, SU(3) helpers, projectSU3, Landau OR and
   microcanonical sweeps, Landau quality -- compared bit for bit.
2. The same comparisons live against oracle/_ref where it is built.
3. Published Random123 known answers for Philox4x32-10.
4. Closed forms and invariants that a dropped term, a wrong sign or a
   transposed operand would break: unitarity / det = 1 of reconstructed and
   projected links, the OR and microcanonical update formulas, monotonicity of
   the functional under omega = 1 OR, exact conservation under microcanonical
   steps, gauge invariance of the plaquette, and the known answer of a pure
   gauge (gff -> 1, theta -> 0).
"""
import numpy as np
import pytest

from oracle_lib import LANDAU, MAG, MICRO, OR, U1XU1, from_matrices, rows01, to_matrices

GOLDEN = None


def golden():
    global GOLDEN
    if GOLDEN is None:
        from pathlib import Path
        GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "ref_vectors.npz")
    return GOLDEN


SHAPES = [(4, 4, 4, 4), (4, 6, 4, 8), (6, 4, 8, 4), (2, 2, 2, 2), (1, 8, 6, 4)]


def _key(dims):
    return "x".join(map(str, dims))


def m3(a18):
    return (np.asarray(a18[0::2]) + 1j * np.asarray(a18[1::2])).reshape(3, 3)


def random_lattice(dims, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(dims))
    z = rng.standard_normal((4 * V, 3, 3)) + 1j * rng.standard_normal((4 * V, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    q = q / np.linalg.det(q)[:, None, None] ** (1 / 3)
    return from_matrices(q.reshape(4, V, 3, 3))


def plaquette(U, dims):
    """Average Re tr of the plaquettes / 3 (gauge invariant)."""
    V = int(np.prod(dims))
    M = to_matrices(U, V)
    c = np.stack(np.unravel_index(np.arange(V), dims), axis=1)

    def pidx(cc):
        lex = np.ravel_multi_index(cc.T, dims)
        return np.where(cc.sum(axis=1) % 2 == 0, lex // 2, lex // 2 + V // 2)

    here = pidx(c)
    tot = 0.0
    for mu in range(4):
        cm = c.copy()
        cm[:, mu] = (cm[:, mu] + 1) % dims[mu]
        xm = pidx(cm)
        for nu in range(mu + 1, 4):
            cn = c.copy()
            cn[:, nu] = (cn[:, nu] + 1) % dims[nu]
            xn = pidx(cn)
            P = M[mu, here] @ M[nu, xm] @ np.conj(np.swapaxes(M[mu, xn], -1, -2)) @ \
                np.conj(np.swapaxes(M[nu, here], -1, -2))
            tot += np.trace(P, axis1=-2, axis2=-1).real.sum()
    return tot / (6 * V * 3)


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("dims", SHAPES)
def test_geometry_golden(oracle, dims):
    g = golden()
    k = _key(dims)
    assert np.array_equal(oracle.neighbour_table(dims), g[f"nn_{k}"])
    V = int(np.prod(dims))
    coords = g[f"coords_{k}"]
    for idx in range(V):
        assert oracle.site_coords(dims, idx) == tuple(coords[idx])
        assert oracle.site_index(dims, coords[idx]) == idx
    for s, mu, i, j, c, expect in g[f"pattern_{k}"]:
        assert oracle.lib.orc_pattern_index(V, int(s), int(mu), int(i), int(j), int(c)) == expect


def test_philox_golden(oracle):
    g = golden()
    for ctr, key, exp in zip(g["philox_ctr"], g["philox_key"], g["philox_out"]):
        assert oracle.philox(tuple(int(x) for x in ctr), tuple(int(x) for x in key)) == tuple(int(x) for x in exp)
    for w, exp in zip(g["u01_64_in"], g["u01_64_out"]):
        assert oracle.lib.orc_u01_64(int(w)) == exp
    for w, exp in zip(g["u01_32_in"], g["u01_32_out"]):
        assert np.float32(oracle.lib.orc_u01_32(int(w))) == exp


def test_philox_published_kat(oracle):
    """Random123 known-answer vectors for philox4x32 with 10 rounds."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, exp in kat:
        assert oracle.philox(ctr, key) == exp


def test_u01_closed_forms(oracle):
    # u01_open_open_64_53: (1 + (w >> 11)) * (2^53 - 1) * 2^-106, strictly inside (0, 1)
    for w in (0, 1, 2**11, 2**63, 2**64 - 1):
        assert oracle.lib.orc_u01_64(w) == (1 + (w >> 11)) * (2.0**53 - 1) * 2.0**-106
    assert 0 < oracle.lib.orc_u01_64(0) < oracle.lib.orc_u01_64(2**64 - 1) < 1
    assert 0 < oracle.lib.orc_u01_32(0) < oracle.lib.orc_u01_32(2**32 - 1) < 1


def test_philox_stream_word_order(oracle):
    """PhiloxWrapper (PhiloxWrapper.hxx:93-140): first block has counter word 0
    = 1, FP64 draws take the 64-bit words [1] then [0] of each block."""
    tid, seed, counter = 1234, 7, 99
    draws = oracle.philox_stream(tid, seed, counter, 6)
    for b in range(3):
        w = oracle.philox((b + 1, counter, 0x12345678, 0xABCDEF09), (tid, seed))
        hi = w[2] | (w[3] << 32)
        lo = w[0] | (w[1] << 32)
        assert draws[2 * b] == oracle.lib.orc_u01_64(hi)
        assert draws[2 * b + 1] == oracle.lib.orc_u01_64(lo)


def test_su3_helpers_golden(oracle):
    g = golden()
    subs = [(0, 2), (1, 2), (0, 1)]
    for n, m in enumerate(g["su3_in"]):
        assert np.array_equal(oracle.mat_op("orc_reconstruct_third_line", m), g["su3_reconstruct"][n])
        assert np.array_equal(oracle.mat_op("orc_project_su3", m), g["su3_project"][n])
        for s, (i, j) in enumerate(subs):
            assert np.array_equal(oracle.subgroup_quaternion(m, i, j), g["su3_quat"][n][s])
            assert np.array_equal(oracle.subgroup_mult("left", m, i, j, g["su3_q"][n]), g["su3_left"][n][s])
            assert np.array_equal(oracle.subgroup_mult("right", m, i, j, g["su3_q"][n]), g["su3_right"][n][s])


def test_lattice_ops_golden(oracle):
    g = golden()
    dims = tuple(int(x) for x in g["lat_dims"])
    U = g["lat_in"].copy()
    assert oracle.quality(LANDAU, dims, U) == tuple(g["lat_quality_in"])
    P = U.copy()
    oracle.project_all(dims, P)
    assert np.array_equal(P, g["lat_project"])
    W = U.copy()
    for _ in range(3):
        for parity in (0, 1):
            oracle.sweep(LANDAU, OR, dims, W, parity, 1.7)
    assert np.array_equal(W, g["lat_or3"])
    assert oracle.quality(LANDAU, dims, W) == tuple(g["lat_quality_or3"])
    for parity in (0, 1):
        oracle.sweep(LANDAU, MICRO, dims, W, parity, 0.0)
    assert np.array_equal(W, g["lat_or3_micro1"])


# --------------------------------------------------- live reference harness
@pytest.mark.parametrize("dims", [(4, 6, 4, 8), (6, 4, 8, 4)])
@pytest.mark.parametrize("algo", [OR, MICRO])
def test_sweep_vs_reference_harness(oracle, ref, dims, algo):
    U = random_lattice(dims, 3)
    A, B = U.copy(), U.copy()
    for _ in range(2):
        for parity in (0, 1):
            oracle.sweep(LANDAU, algo, dims, A, parity, 1.7)
            ref.sweep_landau(algo, dims, B, parity, 1.7)
    assert np.array_equal(A, B)
    assert oracle.quality(LANDAU, dims, A) == ref.quality_landau(dims, B)


def test_geometry_vs_reference_harness(oracle, ref):
    for dims in [(8, 4, 6, 2), (2, 6, 4, 10)]:
        assert np.array_equal(oracle.neighbour_table(dims), ref.neighbour_table(dims))


# ----------------------------------------------------------- closed forms
def test_reconstruct_and_project_are_su3(oracle):
    g = golden()
    for m in g["su3_in"]:
        p = m3(oracle.mat_op("orc_project_su3", m))
        assert np.abs(p @ p.conj().T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(p) - 1) < 1e-14
        # projection of an SU(3) matrix is the identity map
        q = m3(oracle.mat_op("orc_project_su3", oracle.mat_op("orc_project_su3", m)))
        assert np.abs(q - p).max() < 1e-15
        # third row = conj(row0 x row1) for a unitary input
        r = m3(oracle.mat_op("orc_reconstruct_third_line", oracle.mat_op("orc_project_su3", m)))
        assert np.abs(r[2] - np.conj(np.cross(p[0], p[1]))).max() < 1e-15


def test_subgroup_multiplication_closed_form(oracle):
    """left: rows i,j <- G rows i,j ; right: cols i,j <- cols i,j G, with
    G(q) = [[q0 + i q3, q2 + i q1], [-q2 + i q1, q0 - i q3]] (Quaternion.hxx:100-132)."""
    g = golden()
    rng = np.random.default_rng(5)
    for m in g["su3_in"][:3]:
        M = m3(m)
        q = rng.standard_normal(4)
        G = np.array([[q[0] + 1j * q[3], q[2] + 1j * q[1]], [-q[2] + 1j * q[1], q[0] - 1j * q[3]]])
        for i, j in [(0, 2), (1, 2), (0, 1)]:
            L = M.copy()
            L[[i, j], :] = G @ M[[i, j], :]
            assert np.abs(m3(oracle.subgroup_mult("left", m, i, j, q)) - L).max() < 1e-14
            R = M.copy()
            R[:, [i, j]] = M[:, [i, j]] @ G
            assert np.abs(m3(oracle.subgroup_mult("right", m, i, j, q)) - R).max() < 1e-14
            # quaternion of the 2x2 block: (Re a00 + Re a11, Im a01 + Im a10, Re a01 - Re a10, Im a00 - Im a11)
            a = M[np.ix_([i, j], [i, j])]
            expect = [a[0, 0].real + a[1, 1].real, a[0, 1].imag + a[1, 0].imag, a[0, 1].real - a[1, 0].real,
                      a[0, 0].imag - a[1, 1].imag]
            assert np.allclose(oracle.subgroup_quaternion(m, i, j), expect, atol=1e-15)


def test_or_omega1_increases_functional(oracle):
    """With omega = 1 every subgroup step maximises the local Landau functional,
    so gff never decreases (Cabibbo-Marinari); theta falls."""
    dims = (4, 6, 4, 8)
    U = random_lattice(dims, 11)
    g_prev, t0 = oracle.quality(LANDAU, dims, U)
    for _ in range(6):
        for parity in (0, 1):
            oracle.sweep(LANDAU, OR, dims, U, parity, 1.0)
        g, t = oracle.quality(LANDAU, dims, U)
        assert g >= g_prev - 1e-14
        g_prev = g
    assert t < t0


@pytest.mark.parametrize("gauge", [LANDAU, U1XU1])
def test_micro_conserves_functional(oracle, gauge):
    dims = (4, 6, 4, 8)
    U = random_lattice(dims, 12)
    g0, t0 = oracle.quality(LANDAU, dims, U)
    for _ in range(5):
        for parity in (0, 1):
            oracle.sweep(gauge, MICRO, dims, U, parity, 0.0)
    g, t = oracle.quality(LANDAU, dims, U)
    assert g == pytest.approx(g0, abs=1e-13)
    assert t != pytest.approx(t0, rel=1e-3)


@pytest.mark.parametrize("gauge", [LANDAU, MAG, U1XU1])
def test_sweeps_are_gauge_transformations(oracle, gauge):
    """Every sweep multiplies the links at a site by G and G^dagger: the
    plaquette is unchanged."""
    dims = (4, 4, 4, 4)
    U = random_lattice(dims, 13)
    p0 = plaquette(U, dims)
    for _ in range(3):
        for parity in (0, 1):
            oracle.sweep(gauge, OR, dims, U, parity, 1.7)
    oracle.project_all(dims, U)
    assert plaquette(U, dims) == pytest.approx(p0, abs=1e-13)


def test_pure_gauge_converges_to_unit_functional(oracle, gf):
    """Unit links after a random gauge transformation: Landau gauge fixing
    brings the functional back to 1 and theta to 0."""
    dims = (4, 4, 4, 4)
    U = gf.make_config(dims, 0.0, 1, 7)
    g0, t0 = oracle.quality(LANDAU, dims, U)
    assert g0 < 0.5 and t0 > 1
    n, g, t = oracle.run_or(LANDAU, dims, U, 1.7, 1e-14, 3000, 1, 100)
    assert t < 1e-14
    assert g == pytest.approx(1.0, abs=1e-12)


def test_cold_configuration(oracle, gf):
    dims = (4, 6, 4, 8)
    U = gf.make_config(dims, 0.0, 1, 0)
    assert oracle.quality(LANDAU, dims, U) == (1.0, 0.0)
    g, t = oracle.quality(MAG, dims, U)
    assert g == 1.0 and t == 0.0


def test_run_or_loop_semantics(oracle, gf):
    """checkprecision > 1 quantises the stopping sweep: the loop breaks after
    a sweep i with i % checkprecision == 0 (LandauGaugeFixingSU3_4D.cu:254-259)."""
    dims = (4, 4, 4, 4)
    U = gf.make_config(dims, 0.3, 1, 2)
    A = U.copy()
    n1, _, _ = oracle.run_or(LANDAU, dims, A, 1.7, 1e-10, 5000, 1, 100)
    B = U.copy()
    n10, _, t10 = oracle.run_or(LANDAU, dims, B, 1.7, 1e-10, 5000, 10, 100)
    assert (n10 - 1) % 10 == 0 and n10 >= n1 and t10 < 1e-10
    C = U.copy()
    n0, _, _ = oracle.run_or(LANDAU, dims, C, 1.7, 1e-10, 7, 1, 100)
    assert n0 == 7
