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
5. The random update rules (heatbath, random transformation, stochastic
   relaxation) one quaternion at a time: Kolmogorov-Smirnov tests against their
   closed-form sampling densities, with power checks, and exact outcomes.
6. MAG theta against half the gradient of the MAG functional under site
   rotations (exact, from three samples of a trigonometric polynomial), a
   closed form on a cold lattice rotated at one site, and its vanishing at
   a converged maximum.
"""
import numpy as np
import pytest

from oracle_lib import LANDAU, MAG, MICRO, OR, RANDOM, SA, SR, U1XU1, from_matrices, rows01, to_matrices

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


# ------------------------------------------ pins of the non-Landau branches
# cyclic permutation: an SU(3) matrix with zero diagonal, so Re tr P = 0
PERM = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], dtype=complex)


def test_coulomb_quality_is_landau_quality_without_time_links(oracle):
    """The Coulomb quality of a timeslice (CoulombKernelsSU3.hxx:110-183) is
    the Landau quality (pinned by the reference-header golden vectors) with
    the time direction left out.  Embedded in a {2, Nx, Ny, Nz} lattice whose
    time links are all the same permutation matrix P, the time terms of the
    Landau per-site values are exactly zero (P - P in Delta, Re tr P = 0 in
    the functional), so the per-site values agree bit for bit."""
    from oracle_lib import COULOMB
    dims3 = (1, 6, 4, 8)
    V3 = int(np.prod(dims3))
    up = random_lattice(dims3, 21)
    dims4 = (2,) + dims3[1:]
    V4 = int(np.prod(dims4))
    A = to_matrices(up, V3)
    M = np.empty((4, V4, 3, 3), dtype=complex)
    M[0] = PERM
    idx0 = np.array([oracle.site_index(dims4, (0,) + oracle.site_coords(dims3, i)[1:]) for i in range(V3)])
    other = to_matrices(random_lattice(dims3, 23), V3)
    idx1 = np.array([oracle.site_index(dims4, (1,) + oracle.site_coords(dims3, i)[1:]) for i in range(V3)])
    for mu in (1, 2, 3):
        M[mu, idx0] = A[mu]
        M[mu, idx1] = other[mu]
    g3, a3 = oracle.quality_per_site(COULOMB, dims3, up)
    g4, a4 = oracle.quality_per_site(LANDAU, dims4, from_matrices(M))
    assert np.array_equal(g3, g4[idx0]) and np.array_equal(a3, a4[idx0])


def _slice_plaquettes(oracle, dims3, up, dw, nxt):
    """Traces of the spatial plaquettes of slice t (up) and of the temporal
    plaquettes between slices t-1 (dw) and t, and between t and t+1 (nxt)."""
    V3 = int(np.prod(dims3))
    Mu, Md, Mn = to_matrices(up, V3), to_matrices(dw, V3), to_matrices(nxt, V3)
    fw = {}
    for i in range(V3):
        c = oracle.site_coords(dims3, i)
        for mu in (1, 2, 3):
            d = list(c)
            d[mu] = (d[mu] + 1) % dims3[mu]
            fw[(i, mu)] = oracle.site_index(dims3, tuple(d))
    dag = lambda m: np.conj(m.T)
    out = []
    for i in range(V3):
        for mu in (1, 2, 3):
            for nu in range(mu + 1, 4):
                P = Mu[mu, i] @ Mu[nu, fw[(i, mu)]] @ dag(Mu[mu, fw[(i, nu)]]) @ dag(Mu[nu, i])
                out.append(np.trace(P))
            # (t-1, t): U_i(x, t-1) U_0(x+i, t-1) U_i(x, t)^dag U_0(x, t-1)^dag
            P = Md[mu, i] @ Md[0, fw[(i, mu)]] @ dag(Mu[mu, i]) @ dag(Md[0, i])
            out.append(np.trace(P))
            # (t, t+1): U_i(x, t) U_0(x+i, t) U_i(x, t+1)^dag U_0(x, t)^dag
            P = Mu[mu, i] @ Mu[0, fw[(i, mu)]] @ dag(Mn[mu, i]) @ dag(Mu[0, i])
            out.append(np.trace(P))
    return np.array(out)


def test_coulomb_sweeps(oracle):
    """Coulomb sweeps of a timeslice:
    * they are a gauge transformation of slice t that also multiplies the time
      links of slice t from the left and those of slice t-1 from the right
      (CoulombKernelsSU3.hxx:208-250): every spatial plaquette of slice t and
      every temporal plaquette to slices t-1 and t+1 keeps its trace;
    * microcanonical steps conserve the Coulomb (spatial) functional exactly,
      and omega = 1 OR never lowers it -- both fail if the time links entered
      the subgroup sum."""
    from oracle_lib import COULOMB
    dims3 = (1, 6, 4, 8)
    up = random_lattice(dims3, 31)
    dw = random_lattice(dims3, 32)
    nxt = random_lattice(dims3, 33)
    p0 = _slice_plaquettes(oracle, dims3, up, dw, nxt)
    for _ in range(3):
        for parity in (0, 1):
            oracle.sweep_coulomb(OR, dims3, up, dw, parity, 1.7)
    oracle.project_all(dims3, up)
    oracle.project_all(dims3, dw)
    assert np.abs(_slice_plaquettes(oracle, dims3, up, dw, nxt) - p0).max() < 1e-13
    g0, t0 = oracle.quality(COULOMB, dims3, up)
    for _ in range(4):
        for parity in (0, 1):
            oracle.sweep_coulomb(MICRO, dims3, up, dw, parity, 0.0)
    g1, t1 = oracle.quality(COULOMB, dims3, up)
    assert g1 == pytest.approx(g0, abs=1e-13) and t1 != pytest.approx(t0, rel=1e-3)
    g_prev = g1
    for _ in range(4):
        for parity in (0, 1):
            oracle.sweep_coulomb(OR, dims3, up, dw, parity, 1.0)
        g, _ = oracle.quality(COULOMB, dims3, up)
        assert g >= g_prev - 1e-14
        g_prev = g
    assert g_prev > g1 + 1e-3


@pytest.mark.parametrize("algo,param", [(OR, 1.7), (MICRO, 0.0), (OR, 1.0)])
def test_u1xu1_sweeps_are_diagonal_rotations(oracle, algo, param):
    """U1xU1 restricts every subgroup element to (q0, 0, 0, q3), a diagonal
    phase rotation (GaugeFixingSubgroupStep.hxx:238-353): the moduli |U_ij|
    of every link are unchanged."""
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    U = random_lattice(dims, 24)
    oracle.project_all(dims, U)
    m0 = np.abs(to_matrices(U, V))
    for _ in range(3):
        for parity in (0, 1):
            oracle.sweep(U1XU1, algo, dims, U, parity, param)
    oracle.project_all(dims, U)
    assert np.abs(np.abs(to_matrices(U, V)) - m0).max() < 1e-14
    # and the links did change
    assert np.abs(to_matrices(U, V) - to_matrices(random_lattice(dims, 24), V)).max() > 1e-3


@pytest.mark.parametrize("gauge", [U1XU1, MAG])
def test_omega1_or_maximises_each_functional(oracle, gauge):
    """With omega = 1 the U1xU1 step maximises the Landau functional over
    diagonal rotations and the MAG step, g0 = D + sqrt(D^2 + E^2 + F^2), the
    MAG functional sum |U_ii|^2 over the subgroup
    (GaugeFixingSubgroupStep.hxx:529-575): neither functional decreases."""
    dims = (4, 6, 4, 8)
    U = random_lattice(dims, 25)
    qual = MAG if gauge == MAG else LANDAU
    g_prev, _ = oracle.quality(qual, dims, U)
    g_first = g_prev
    for _ in range(6):
        for parity in (0, 1):
            oracle.sweep(gauge, OR, dims, U, parity, 1.0)
        g, _ = oracle.quality(qual, dims, U)
        assert g >= g_prev - 1e-14
        g_prev = g
    assert g_prev > g_first + 1e-3


def _local_max_check(oracle, gauge, functional, generators):
    """After an omega = 1 OR half-sweep every updated site's last subgroup
    step, (0,1), left the local functional at a maximum over that subgroup:
    small extra rotations exp(i eps sigma_k) in (0,1) -- applied to the up
    links from the left and to the down links from the right -- do not raise
    it.  A wrong term, sign or factor in the subgroup sum moves the update off
    the maximum and some rotation raises the functional at first order."""
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    U = random_lattice(dims, 41)
    oracle.sweep(gauge, OR, dims, U, 0, 1.0)
    oracle.project_all(dims, U)
    M = to_matrices(U, V)
    nn = oracle.neighbour_table(dims).reshape(8, V)
    even = np.arange(V // 2)
    up = M[:, even]                                        # [mu, site, 3, 3]
    dn = np.stack([M[mu, nn[2 * mu, even]] for mu in range(4)])
    base = functional(up, dn)
    sig = {1: np.array([[0, 1], [1, 0]]), 2: np.array([[0, -1j], [1j, 0]]), 3: np.array([[1, 0], [0, -1]])}
    worst = -np.inf
    for k in generators:
        for eps in (1e-3, -1e-3):
            R = np.eye(3, dtype=complex)
            R[:2, :2] = np.cos(eps) * np.eye(2) + 1j * np.sin(eps) * sig[k]
            f = functional(R @ up, dn @ np.conj(R.T))
            worst = max(worst, (f - base).max())
    assert worst < 1e-12, worst


def test_mag_update_is_local_maximum(oracle):
    """MAG (GaugeFixingSubgroupStep.hxx:354-575): the functional is the sum of
    |U_ii|^2 over the site's eight links; the subgroup step excludes sigma_3
    (q3 = 0), which leaves it unchanged."""
    def f(up, dn):
        d = np.abs(np.diagonal(up, axis1=-2, axis2=-1)) ** 2 + np.abs(np.diagonal(dn, axis1=-2, axis2=-1)) ** 2
        return d.sum(axis=(0, 2))
    _local_max_check(oracle, MAG, f, (1, 2))


def test_u1xu1_update_is_local_maximum(oracle):
    """U1xU1: the Landau functional Re tr over the eight links, maximised over
    the diagonal rotations (sigma_3) alone."""
    def f(up, dn):
        return (np.trace(up, axis1=-2, axis2=-1).real + np.trace(dn, axis1=-2, axis2=-1).real).sum(axis=0)
    _local_max_check(oracle, U1XU1, f, (3,))


def test_landau_update_is_local_maximum(oracle):
    """Landau: the same functional over all three generators."""
    def f(up, dn):
        return (np.trace(up, axis1=-2, axis2=-1).real + np.trace(dn, axis1=-2, axis2=-1).real).sum(axis=0)
    _local_max_check(oracle, LANDAU, f, (1, 2, 3))


# ---------------------------------------------------------------- update rules
# Each rule acts on the subgroup quaternion a accumulated for a site; the local
# functional of a unit quaternion g is the 4D dot product g.a, maximised by
# g = a/|a| (the omega = 1 OR result).  The random rules are pinned through
# their sampling distributions (Kolmogorov-Smirnov against the closed-form
# densities) with a power check that a wrong parameter is rejected.

def _hamilton(p, q):
    p0, p1, p2, p3 = p
    q0, q1, q2, q3 = q
    return np.array([p0 * q0 - p1 * q1 - p2 * q2 - p3 * q3,
                     p0 * q1 + p1 * q0 + p2 * q3 - p3 * q2,
                     p0 * q2 - p1 * q3 + p2 * q0 + p3 * q1,
                     p0 * q3 + p1 * q2 - p2 * q1 + p3 * q0])


def _rand_quats(n, seed, scale=2.0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 4)) * scale


def _samples(oracle, algo, a, param, n, seed=11, counter=5):
    return np.array([oracle.update(algo, a, param, tid, seed, counter) for tid in range(n)])


def _ks_pvalue(x, cdf):
    from scipy import stats
    return stats.kstest(x, cdf).pvalue


def _kp_cdf(k):
    """CDF of x on [-1, 1] with density proportional to sqrt(1 - x^2) exp(k x):
    the SU(2) heatbath distribution of g.a/|a| at beta |a| = k."""
    from scipy.integrate import cumulative_trapezoid
    xs = np.linspace(-1.0, 1.0, 400001)
    w = np.sqrt(np.clip(1.0 - xs * xs, 0.0, None)) * np.exp(k * (xs - 1.0))
    c = cumulative_trapezoid(w, xs, initial=0.0)
    c /= c[-1]
    return lambda x: np.interp(x, xs, c)


def test_micro_update_is_reflection_through_optimum(oracle):
    """MicroUpdate.hxx:43-66: g' = (a/|a|)^2, the reflection of g = 1 through the
    maximiser; it keeps the local functional (g'.a = a0) and is a unit quaternion."""
    for a in _rand_quats(200, 3):
        g = oracle.update(MICRO, a)
        ah = a / np.linalg.norm(a)
        np.testing.assert_allclose(g, _hamilton(ah, ah), atol=1e-14)
        assert abs(g @ a - a[0]) < 1e-13 * np.linalg.norm(a)


def test_or_update_closed_forms(oracle):
    """OrUpdate.hxx:52-78: unit result whose vector part is parallel to a's;
    omega = 1 gives a/|a|; to first order in the angle the rotation angle of
    a/|a| is multiplied by omega (the overrelaxation g -> g^omega)."""
    for a in _rand_quats(100, 4):
        np.testing.assert_allclose(oracle.update(OR, a, 1.0), a / np.linalg.norm(a), atol=1e-15)
        for w in (1.3, 1.7, 1.95):
            g = oracle.update(OR, a, w)
            assert abs(np.linalg.norm(g) - 1.0) < 1e-14
            v, u = g[1:], a[1:]
            np.testing.assert_allclose(np.cross(v, u), 0.0, atol=1e-13)
            assert v @ u > 0 and np.sign(g[0]) == np.sign(a[0])
    for w in (1.3, 1.7, 1.95):
        for th in (1e-3, 3e-3):
            n = np.array([0.3, -0.5, 0.8]); n /= np.linalg.norm(n)
            a = 1.7 * np.concatenate([[np.cos(th)], np.sin(th) * n])
            g = oracle.update(OR, a, w)
            ang = np.arctan2(np.linalg.norm(g[1:]), g[0])
            assert abs(ang / th - w) < 2 * th, (w, th, ang / th)


@pytest.mark.parametrize("norm,temperature", [(2.0, 4.0), (2.0, 1.0), (3.0, 0.375)])
def test_sa_update_is_heatbath(oracle, norm, temperature):
    """SaUpdate.hxx:52-147 (Kennedy-Pendleton): g.a/|a| follows
    sqrt(1 - x^2) exp(|a| x / T), the component of g orthogonal to a/|a| points
    uniformly over the 2-sphere, and g is a unit quaternion.  The key (tid) runs
    over 0..n-1, one independent stream per sample."""
    n = 12000
    a = np.array([0.4, -0.7, 0.5, 0.3])
    a = a / np.linalg.norm(a) * norm
    ah = a / norm
    g = _samples(oracle, SA, a, temperature, n)
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-13)
    x = g @ ah
    k = norm / temperature
    assert _ks_pvalue(x, _kp_cdf(k)) > 1e-3
    # power: a 50% error in beta |a| is rejected
    assert _ks_pvalue(x, _kp_cdf(1.5 * k)) < 1e-6
    assert _ks_pvalue(x, _kp_cdf(k / 1.5)) < 1e-6
    # isotropy of the orthogonal part
    basis = np.linalg.qr(np.column_stack([ah, np.eye(4)[:, :3]]))[0][:, 1:]
    v = g @ basis
    u = v / np.linalg.norm(v, axis=1, keepdims=True)
    assert np.abs(u.mean(axis=0)).max() < 5 / np.sqrt(3 * n)
    np.testing.assert_allclose(u.T @ u / n, np.eye(3) / 3, atol=5 * np.sqrt(2 / 15 / n))


def test_random_update_distribution(oracle):
    """RandomUpdate.hxx:47-59: the result does not depend on a; a0 = cos(pi alpha)
    with alpha uniform on [0, 1), and the vector part's direction is uniform on
    the 2-sphere (cos theta uniform on [-1, 1], phi uniform on [0, 2 pi))."""
    from scipy import stats
    n = 12000
    g = _samples(oracle, RANDOM, np.array([1.0, 0.2, -0.3, 0.4]), 0.0, n)
    g2 = _samples(oracle, RANDOM, np.array([-0.1, 2.0, 0.5, -0.6]), 0.0, 50)
    np.testing.assert_array_equal(g[:50], g2)
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-14)
    alpha = np.arccos(np.clip(g[:, 0], -1, 1)) / np.pi
    assert stats.kstest(alpha, "uniform").pvalue > 1e-3
    assert stats.kstest(alpha ** 1.1, "uniform").pvalue < 1e-6
    s = np.linalg.norm(g[:, 1:], axis=1)
    ct = g[:, 3] / s
    assert stats.kstest(ct, stats.uniform(-1, 2).cdf).pvalue > 1e-3
    phi = np.mod(np.arctan2(g[:, 2], g[:, 1]), 2 * np.pi) / (2 * np.pi)
    assert stats.kstest(phi, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("p", [0.0, 0.3, 1.0])
def test_sr_update_flip_frequency(oracle, p):
    """SrUpdate.hxx:54-96 (stochastic relaxation): with probability p the result
    is the normalised square a^2/|a|^2 (the overrelaxed step), otherwise
    a/|a|; the first draw of the thread's stream decides."""
    n = 6000
    quats = _rand_quats(n, 9)
    flips = 0
    for tid, a in enumerate(quats):
        g = oracle.update(SR, a, p, tid, 3, 17)
        ah = a / np.linalg.norm(a)
        sq = _hamilton(ah, ah)
        if np.abs(g - sq).max() < 1e-13:
            flips += 1
        else:
            np.testing.assert_allclose(g, ah, atol=1e-14)
    if p in (0.0, 1.0):
        assert flips == int(p) * n
    else:
        assert abs(flips / n - p) < 5 * np.sqrt(p * (1 - p) / n), flips / n


# ------------------------------------------------------------ MAG theta
# MAGKernelsSU3.hxx:116-223: theta_x = sum_k |X_k| over the subgroups
# k = (0,1), (0,2), (1,2), with X_k summed from the up links' 2x2 blocks and
# the down links' Hermitian blocks.  The pins below use only the MAG
# functional, sum_j |U_jj|^2 over a site's eight links, and rotations: for a
# gauge rotation exp(i t sigma) of site x in subgroup k (up links rotated from
# the left, down links from the right by its inverse), the functional is
# A + B cos 2t + C sin 2t, and X_k is half its gradient at t = 0 over
# sigma_2 (real part) and sigma_1 (imaginary part).  With that trigonometric
# form the derivative is exactly F(pi/4) - F(-pi/4): no finite-difference
# error.  A squared |X_k|, a dropped or non-Hermitian down term, a wrong slot
# i+j-1 or a factor slip each move theta_x off half the gradient norm.
MAG_SUBGROUPS = ((0, 1), (0, 2), (1, 2))
PAULI = {1: np.array([[0, 1], [1, 0]], dtype=complex), 2: np.array([[0, -1j], [1j, 0]])}


def _su2_embed(i, j, k, t):
    R = np.eye(3, dtype=complex)
    R[np.ix_([i, j], [i, j])] = np.cos(t) * np.eye(2) + 1j * np.sin(t) * PAULI[k]
    return R


def _site_links(oracle, dims, U):
    """Up links U_mu(x) and down links U_mu(x - mu) of every site: [mu, idx, 3, 3]."""
    V = int(np.prod(dims))
    M = to_matrices(U, V)
    nn = oracle.neighbour_table(dims).reshape(8, V)
    return M, np.stack([M[mu, nn[2 * mu]] for mu in range(4)])


def _mag_local(up, dn):
    d = np.abs(np.diagonal(up, axis1=-2, axis2=-1)) ** 2 + np.abs(np.diagonal(dn, axis1=-2, axis2=-1)) ** 2
    return d.sum(axis=(0, 2))


def _mag_theta_from_gradient(up, dn):
    theta = 0.0
    for i, j in MAG_SUBGROUPS:
        grad = []
        for k in (1, 2):
            fp, fm = [_mag_local(R @ up, dn @ np.conj(R.T)) for R in
                      (_su2_embed(i, j, k, np.pi / 4), _su2_embed(i, j, k, -np.pi / 4))]
            grad.append(fp - fm)
        theta = theta + 0.5 * np.hypot(grad[0], grad[1])
    return theta


@pytest.mark.parametrize("kind", ["hot", "smooth"])
def test_mag_theta_is_half_the_gauge_gradient(oracle, gf, kind):
    dims = (4, 6, 4, 8)
    U = random_lattice(dims, 13) if kind == "hot" else gf.make_config(dims, 0.35, 1, 2)
    up, dn = _site_links(oracle, dims, U)
    gff_x, theta_x = oracle.quality_per_site(MAG, dims, U)
    expect = _mag_theta_from_gradient(up, dn)
    assert expect.min() > 1e-3  # far from a maximum: every site has a gradient
    np.testing.assert_allclose(theta_x, expect, rtol=1e-12, atol=1e-13)
    # gff_x: sum_j |U_mu(x)_jj|^2 over the up links alone (MAGKernelsSU3.hxx:184-186)
    np.testing.assert_allclose(gff_x, (np.abs(np.diagonal(up, axis1=-2, axis2=-1)) ** 2).sum(axis=(0, 2)),
                               rtol=1e-14)


@pytest.mark.parametrize("sub", MAG_SUBGROUPS)
@pytest.mark.parametrize("k", [1, 2])
def test_mag_theta_gauge_rotation_closed_form(oracle, sub, k):
    """A cold lattice rotated at one site x0 by g = exp(i eps sigma_k) in one
    subgroup: the eight links of x0 contribute -i sin 2 eps (sigma_1) or
    -sin 2 eps (sigma_2) each, so theta_x0 = 8 |sin 2 eps|; each of the eight
    neighbours sees one rotated link, |sin 2 eps|; every other site 0.  The
    lattice values, with the reference's prefactors 1/Nc and 1/(Nc Ndim)
    (MAGKernelsSU3.hxx:76-83) and the average over V: theta = 16 sin 2 eps /
    (3 V), gff = 1 - 4 sin^2 eps / (3 V)."""
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    eps = 0.3
    g = _su2_embed(*sub, k, eps)
    M = np.broadcast_to(np.eye(3, dtype=complex), (4, V, 3, 3)).copy()
    nn = oracle.neighbour_table(dims).reshape(8, V)
    x0 = 37
    for mu in range(4):
        M[mu, x0] = g
        M[mu, nn[2 * mu, x0]] = np.conj(g.T)
    U = from_matrices(M)
    _, theta_x = oracle.quality_per_site(MAG, dims, U)
    want = np.zeros(V)
    want[x0] = 8 * np.sin(2 * eps)
    for mu in range(4):
        want[nn[2 * mu, x0]] = np.sin(2 * eps)
        want[nn[2 * mu + 1, x0]] = np.sin(2 * eps)
    np.testing.assert_allclose(theta_x, want, atol=1e-15, rtol=1e-14)
    gff, theta = oracle.quality(MAG, dims, U)
    assert theta == pytest.approx(16 * np.sin(2 * eps) / (3 * V), rel=1e-14)
    assert gff == pytest.approx(1 - 4 * np.sin(eps) ** 2 / (3 * V), rel=1e-14)


def test_mag_theta_vanishes_at_converged_maximum(oracle, gf):
    """OR run by the oracle to theta < 1e-14 ends at a maximum of the MAG
    functional: no site rotation in any subgroup raises it at first order,
    and every theta_x is at rounding level.  A down term with the wrong sign
    would not cancel the up terms there."""
    dims = (2, 4, 2, 4)
    U = gf.make_config(dims, 0.3, 1, 2)
    n, g, t = oracle.run_or(MAG, dims, U, 1.7, 1e-14, 5000, 1, 100)
    assert n < 5000 and t < 1e-14
    up, dn = _site_links(oracle, dims, U)
    assert np.abs(_mag_theta_from_gradient(up, dn)).max() < 1e-13
    _, theta_x = oracle.quality_per_site(MAG, dims, U)
    assert theta_x.max() < 1e-13
    base = _mag_local(up, dn)
    for i, j in MAG_SUBGROUPS:
        for k in (1, 2):
            for eps in (1e-3, -1e-3):
                R = _su2_embed(i, j, k, eps)
                assert (_mag_local(R @ up, dn @ np.conj(R.T)) - base).max() < 1e-12


# ------------------------------------------------ double-precision updates
def _exact_square_quats(n, seed):
    """Float quaternions whose components are multiples of 2^-6 in (-2, 2):
    their squares and the sums of three squares are exact in float, so the
    reference's float-typed squares (shA is Real) lose nothing and a
    double-precision update is the float rounding of the exact result."""
    rng = np.random.default_rng(seed)
    q = rng.integers(-127, 128, size=(n, 4)).astype(np.float64) / 64.0
    q[np.abs(q).sum(axis=1) == 0] = 1.0
    return q


def test_dp_updates_are_rounded_double_results(oracle32, oracle32dp):
    """DPUPDATES (OrUpdate.hxx:54-64, MicroUpdate.hxx:45-54): on float links
    the update is computed in double and rounded once, so on inputs with exact
    squares OR with omega = 1 gives float(a/|a|) and the microcanonical update
    float((a/|a|)^2) exactly (closed forms, computed here in float64 without
    the update's formula); the float build misses these by an ulp or more on
    some inputs, so the test tells the two precisions apart."""
    dp_or = dp_mi = sp_or = sp_mi = 0
    for a in _exact_square_quats(400, 21):
        ah = a / np.linalg.norm(a)
        want_or = ah.astype(np.float32)
        want_mi = _hamilton(ah, ah).astype(np.float32)
        dp_or += not np.array_equal(oracle32dp.update(OR, a, 1.0), want_or)
        dp_mi += not np.array_equal(oracle32dp.update(MICRO, a), want_mi)
        sp_or += not np.array_equal(oracle32.update(OR, a, 1.0), want_or)
        sp_mi += not np.array_equal(oracle32.update(MICRO, a), want_mi)
    assert dp_or == 0 and dp_mi == 0, (dp_or, dp_mi)
    assert sp_or > 20 and sp_mi > 20, (sp_or, sp_mi)


def test_dp_or_update_follows_reference_typing(oracle32dp):
    """OrUpdate.hxx:54-64 on general float input: the squares are float sums
    (the reference assigns a float expression to a double), b, c and the
    products double, each result rounded to float; omega = 1.7 keeps the
    vector part parallel to a's and the result a unit quaternion to float
    precision."""
    rng = np.random.default_rng(5)
    for a in rng.normal(size=(300, 4)).astype(np.float32):
        ai = np.float64(a[1] * a[1] + a[2] * a[2] + a[3] * a[3])
        a0 = np.float64(a[0] * a[0])
        w = np.float64(np.float32(1.7))
        b = (w * a0 + ai) / (a0 + ai)
        c = 1.0 / np.sqrt(a0 + b * b * ai)
        want = np.array([np.float64(a[0]) * c] + [np.float64(x) * (b * c) for x in a[1:]]).astype(np.float32)
        g = oracle32dp.update(OR, a, 1.7)
        assert np.array_equal(g, want), (a, g, want)
        assert abs(np.linalg.norm(g.astype(np.float64)) - 1.0) < 3e-7
        np.testing.assert_allclose(np.cross(g[1:].astype(np.float64), a[1:].astype(np.float64)), 0.0, atol=1e-6)


@pytest.mark.parametrize("norm,temperature", [(2.0, 1.0), (3.0, 0.375)])
def test_dp_sa_update_is_heatbath(oracle32dp, norm, temperature):
    """SaUpdate.hxx:54-98 (USE_DP_SAUPDATE): the heatbath in double on float
    uniforms has the Kennedy-Pendleton distribution, and its results are unit
    quaternions to float precision."""
    n = 8000
    a = np.array([0.4, -0.7, 0.5, 0.3])
    a = (a / np.linalg.norm(a) * norm).astype(np.float32)
    ah = a.astype(np.float64) / np.linalg.norm(a.astype(np.float64))
    g = _samples(oracle32dp, SA, a, temperature, n).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=3e-7)
    x = g @ ah
    k = np.linalg.norm(a.astype(np.float64)) / temperature
    assert _ks_pvalue(x, _kp_cdf(k)) > 1e-3
    assert _ks_pvalue(x, _kp_cdf(1.5 * k)) < 1e-6
