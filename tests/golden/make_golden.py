"""Write tests/golden/ref_vectors.npz from the reference-header harness.

The harness (oracle/_ref/libsu3gf_ref.so, built by `make -C oracle ref`
from the reference's own lattice and Random123 headers) computes:

* neighbour tables, site index <-> coordinates and GpuPattern indices
  (lattice/SiteIndex.hxx, SiteCoord.hxx, access_pattern/GpuPattern.hxx);
* Philox4x32-10 blocks and u01 conversions (external/Random123/philox.h, u01.h);
* SU(3) helpers on fixed matrices (lattice/SU3.hxx, Quaternion.hxx);
* projectSU3 over a lattice, three FP64 Landau OR sweeps and one
  microcanonical sweep, and the Landau quality, composed from the compiled
  reference classes (oracle/ref_harness.cpp).

The inputs are generated here with numpy (seeded) and stored in the file next
to the outputs, so the fixture stands on its own.  Run from the repository
root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from oracle_lib import RefHarness, from_matrices  # noqa: E402

SHAPES = [(4, 4, 4, 4), (4, 6, 4, 8), (6, 4, 8, 4), (2, 2, 2, 2), (1, 8, 6, 4)]
SWEEP_DIMS = (2, 4, 2, 4)


def random_su3(rng: np.random.Generator, n: int) -> np.ndarray:
    """n Haar-distributed SU(3) matrices (QR of complex Gaussians, phase and det fixed)."""
    z = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / det[:, None, None] ** (1.0 / 3.0)


def to18(m: np.ndarray) -> np.ndarray:
    out = np.empty(18)
    out[0::2] = m.reshape(-1).real
    out[1::2] = m.reshape(-1).imag
    return out


def main() -> None:
    if not RefHarness.available():
        sys.exit("oracle/_ref/libsu3gf_ref.so not built: run `make -C oracle ref` first")
    ref = RefHarness()
    rng = np.random.default_rng(20240517)
    out: dict[str, np.ndarray] = {}

    # geometry
    for dims in SHAPES:
        key = "x".join(map(str, dims))
        V = int(np.prod(dims))
        out[f"nn_{key}"] = ref.neighbour_table(dims)
        coords = np.array([ref.site_coords(dims, i) for i in range(V)], dtype=np.int32)
        out[f"coords_{key}"] = coords
        samp = rng.integers(0, V, size=16)
        pat = [[int(s), mu, i, j, c, ref.pattern_index(dims, int(s), mu, i, j, c)]
               for s in samp for (mu, i, j, c) in [(int(rng.integers(4)), int(rng.integers(3)), int(rng.integers(3)),
                                                    int(rng.integers(2)))]]
        out[f"pattern_{key}"] = np.array(pat, dtype=np.int64)

    # Philox4x32-10 and u01
    ctrs = rng.integers(0, 2**32, size=(16, 4), dtype=np.uint64).astype(np.uint32)
    keys = rng.integers(0, 2**32, size=(16, 2), dtype=np.uint64).astype(np.uint32)
    out["philox_ctr"] = ctrs
    out["philox_key"] = keys
    out["philox_out"] = np.array([ref.philox(tuple(int(x) for x in c), tuple(int(x) for x in k))
                                  for c, k in zip(ctrs, keys)], dtype=np.uint32)
    w64 = np.concatenate([np.array([0, 1, 2**11 - 1, 2**11, 2**63, 2**64 - 1], dtype=np.uint64),
                          rng.integers(0, 2**63, size=10, dtype=np.uint64) * 2])
    w32 = np.concatenate([np.array([0, 1, 255, 256, 2**31, 2**32 - 1], dtype=np.uint64),
                          rng.integers(0, 2**32, size=10, dtype=np.uint64)]).astype(np.uint32)
    out["u01_64_in"] = w64
    out["u01_64_out"] = np.array([ref.lib.ref_u01_64(int(w)) for w in w64])
    out["u01_32_in"] = w32
    out["u01_32_out"] = np.array([ref.lib.ref_u01_32(int(w)) for w in w32], dtype=np.float32)

    # SU(3) helpers on perturbed group elements (rows 0,1 slightly off the group)
    mats = random_su3(rng, 8)
    mats = mats + 1e-3 * (rng.standard_normal(mats.shape) + 1j * rng.standard_normal(mats.shape))
    m18 = np.stack([to18(m) for m in mats])
    out["su3_in"] = m18
    out["su3_reconstruct"] = np.stack([ref.mat_op("ref_reconstruct_third_line", m) for m in m18])
    out["su3_project"] = np.stack([ref.mat_op("ref_project_su3", m) for m in m18])
    subs = [(0, 2), (1, 2), (0, 1)]
    out["su3_quat"] = np.stack([[ref.subgroup_quaternion(m, i, j) for (i, j) in subs] for m in m18])
    qs = rng.standard_normal((8, 4))
    out["su3_q"] = qs
    out["su3_left"] = np.stack([[ref.subgroup_mult("left", m, i, j, q) for (i, j) in subs] for m, q in zip(m18, qs)])
    out["su3_right"] = np.stack([[ref.subgroup_mult("right", m, i, j, q) for (i, j) in subs]
                                 for m, q in zip(m18, qs)])

    # lattice operations on a 2x4x2x4 configuration of random SU(3) links
    dims = SWEEP_DIMS
    V = int(np.prod(dims))
    U = from_matrices(random_su3(rng, 4 * V).reshape(4, V, 3, 3))
    out["lat_dims"] = np.array(dims, dtype=np.int32)
    out["lat_in"] = U
    g0, t0 = ref.quality_landau(dims, U.copy())
    out["lat_quality_in"] = np.array([g0, t0])
    P = U.copy()
    ref.project_all(dims, P)
    out["lat_project"] = P
    W = U.copy()
    for _ in range(3):
        for parity in (0, 1):
            ref.sweep_landau(0, dims, W, parity, 1.7)
    out["lat_or3"] = W.copy()
    out["lat_quality_or3"] = np.array(ref.quality_landau(dims, W.copy()))
    for parity in (0, 1):
        ref.sweep_landau(1, dims, W, parity, 0.0)
    out["lat_or3_micro1"] = W.copy()

    np.savez_compressed(HERE / "ref_vectors.npz", **out)
    print(f"wrote {HERE / 'ref_vectors.npz'} ({(HERE / 'ref_vectors.npz').stat().st_size} bytes)")


if __name__ == "__main__":
    main()
