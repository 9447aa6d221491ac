"""MI355X (gfx950) SU(3) gauge fixing -- Python host side of the C ABI.

The classes below keep the names, argument order and meaning of the
reference's C++ API for the gauge-fixing sweep (paths relative to its src/):

* ``LandauKernelsSU3``   gaugefixing/LandauKernelsSU3.hxx:55-105
* ``CoulombKernelsSU3``  gaugefixing/CoulombKernelsSU3.hxx:53-101
* ``MAGKernelsSU3``      gaugefixing/MAGKernelsSU3.hxx:57-110
* ``U1xU1KernelsSU3``    gaugefixing/U1xU1KernelsSU3.hxx:55-105
* ``CommonKernelsSU3``   gaugefixing/CommonKernelsSU3.hxx:40-62
* ``GaugeFixingStats``   gaugefixing/GaugeFixingStats.hxx:55-81
* ``PhiloxWrapper``      lattice/rng/PhiloxWrapper.hxx:62-150 (host counter)

The reference fixes the lattice extents at compile time; here a ``Lattice``
object carries them.  Device buffers are ``DeviceBuffer`` objects owned by
the library's allocator (no torch: the product path talks to the HIP runtime
only through libsu3gf.so).  Every call goes to the gfx950 kernels; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (  # noqa: F401  (re-exported enums)
    COULOMB,
    CRIT_AVERAGE,
    CRIT_MAX,
    F32,
    F64,
    LANDAU,
    MAG,
    MICRO,
    MIXED_GAUGE32,
    MIXED_GAUGE64,
    MIXED_GAUGE24,
    MIXED_FP64,
    MIXED_LINKS,
    OR,
    PATTERN18,
    R12,
    RANDOM,
    SA,
    SR,
    U1XU1,
    UPDATE_DP,
    SA_GRAPH,
    SA_ONE_LAUNCH,
    WAVEFRONT_PIPELINED,
    FILE_PLAIN,
    FILE_HEADERONLY,
    FILE_VOGT,
    FILE_ILDG,
    FILE_QCDSTAG,
    REAL_STANDARD,
    REAL_DOUBLE,
    REAL_FLOAT,
    PATTERN_GPU,
    PATTERN_GPU_TIMESLICE,
    GFError,
)

__all__ = [
    "Lattice", "DeviceBuffer", "LandauKernelsSU3", "CoulombKernelsSU3", "MAGKernelsSU3",
    "U1xU1KernelsSU3", "CommonKernelsSU3", "GaugeFixingStats", "PhiloxWrapper", "GFError",
    "make_config", "run_or", "run_fast", "run_fast_noalloc", "run_mixed", "run_sa", "batch_run_mixed", "batch_run_sa", "set_batch_launch_blocks",
    "device_info",
    "RunResult", "LinkFile",
]


def _ptr(x) -> C.c_void_p:
    if x is None:
        return C.c_void_p(None)
    if isinstance(x, DeviceBuffer):
        return C.c_void_p(x.ptr)
    if isinstance(x, C.c_void_p):
        return x
    return C.c_void_p(int(x))


# ------------------------------------------------------------------ lattice
class Lattice:
    """Extents {Nt, Nx, Ny, Nz} (mu = 0 is time), precision and link layout.

    ``layout`` 18 is the reference's GpuPattern array (18 reals per link),
    12 the compact array holding rows 0,1 only.  A Coulomb timeslice is a
    lattice with ``dims = (1, Nx, Ny, Nz)``.
    """

    def __init__(self, dims, dtype: int = F64, layout: int = PATTERN18):
        dims = tuple(int(d) for d in dims)
        if len(dims) != 4:
            raise ValueError("dims must be (Nt, Nx, Ny, Nz)")
        self.dims = dims
        self.dtype = dtype
        self.layout = layout
        self.V = dims[0] * dims[1] * dims[2] * dims[3]
        self.Vh = self.V // 2
        self.struct = _lib.gf_lattice((C.c_int * 4)(*dims), dtype, layout)

    @property
    def real_size(self) -> int:
        return 8 if self.dtype == F64 else 4

    @property
    def np_dtype(self):
        return np.float64 if self.dtype == F64 else np.float32

    @property
    def count(self) -> int:
        """Number of reals in the link array."""
        return self.V * 4 * self.layout

    @property
    def nbytes(self) -> int:
        return self.count * self.real_size

    def timeslice(self) -> "Lattice":
        return Lattice((1,) + self.dims[1:], self.dtype, self.layout)

    def with_(self, dtype=None, layout=None) -> "Lattice":
        return Lattice(self.dims, self.dtype if dtype is None else dtype, self.layout if layout is None else layout)

    def __repr__(self) -> str:
        return f"Lattice(dims={self.dims}, dtype={'f64' if self.dtype == F64 else 'f32'}, layout={self.layout})"


# ------------------------------------------------------------------ memory
class DeviceBuffer:
    """Device allocation made through the library (gf_malloc / gf_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _lib.call("gf_malloc", C.byref(p), C.c_size_t(max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        _lib.call("gf_memcpy_h2d", C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes))

    def download(self, dtype=np.float64, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        _lib.call("gf_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer") -> None:
        _lib.call("gf_memcpy_d2d", C.c_void_p(self.ptr), C.c_void_p(other.ptr), C.c_size_t(min(self.nbytes, other.nbytes)))

    def free(self) -> None:
        if self.ptr:
            _lib.load().gf_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def device_info() -> tuple[int, str]:
    n = C.c_int(0)
    name = C.create_string_buffer(128)
    _lib.call("gf_device_info", C.byref(n), name, C.c_size_t(128))
    return n.value, name.value.decode()


def synchronize() -> None:
    _lib.call("gf_synchronize")


# ------------------------------------------------------------------- RNG
class PhiloxWrapper:
    """Host side of the reference's process-wide launch counter
    (PhiloxWrapper.hxx:90,142-150): every RNG kernel launch takes the next value."""

    globalCounter = 0

    @classmethod
    def getNextCounter(cls) -> int:
        c = cls.globalCounter
        cls.globalCounter += 1
        return c

    @classmethod
    def getCurrentCounter(cls) -> int:
        return cls.globalCounter


# ---------------------------------------------------------------- kernels
class _GaugeKernels:
    """Static-wrapper API of one gauge (e.g. LandauKernelsSU3.hxx:55-105).

    ``a`` and ``b`` are the reference's launch dimensions (numBlocks =
    V/2/NSB, threadsPerBlock = 8*NSB, LandauGaugeFixingSU3_4D.cu:124-125);
    they are checked, and NSB = b/8 enters the RNG key exactly as in the
    reference.  ``nnt`` (the neighbour table) is accepted and not read: the
    kernels compute neighbours on the fly with the same formula.
    """

    gauge = LANDAU
    Ndim = 4
    Nc = 3

    def __init__(self, lattice: Lattice, stream=None, dp_updates: bool = False):
        """dp_updates: on FP32 links, compute the OR, microcanonical and
        heatbath updates in double, as the reference's single-precision build
        with DPUPDATES=true (USE_DP_ORUPDATE, USE_DP_MICROUPDATE,
        USE_DP_SAUPDATE; gaugefixing/apps/Makefile:46-48)."""
        self.lattice = lattice
        self.stream = stream
        self.dp_updates = dp_updates

    def initCacheConfig(self) -> None:
        """No-op: the gfx950 kernels need no cache configuration."""

    def _nsb(self, a: int, b: int, V: int) -> int:
        if b <= 0 or b % 8:
            raise GFError(_lib.GF_ERR_BAD_ARG, f"threads per block {b} is not a positive multiple of 8")
        nsb = b // 8
        if a * nsb != V // 2:
            raise GFError(_lib.GF_ERR_BAD_SHAPE, f"launch {a}x{b} does not cover V/2 = {V // 2} sites")
        return nsb

    def _sweep(self, a, b, U, Udw, parity, algo, param=0.0, seed=0, counter=0):
        lat = self.lattice
        nsb = self._nsb(a, b, lat.V)
        if self.dp_updates and algo in (OR, MICRO, SA):
            algo |= UPDATE_DP
        _lib.call("gf_sweep", C.byref(lat.struct), _ptr(U), _ptr(Udw), int(bool(parity)), self.gauge, algo,
                  C.c_float(param), int(seed), C.c_uint32(counter & 0xFFFFFFFF), nsb, _ptr(self.stream))

    def orStep(self, a, b, U, nnt, parity, orParameter):
        self._sweep(a, b, U, None, parity, OR, orParameter)

    def microStep(self, a, b, U, nnt, parity):
        self._sweep(a, b, U, None, parity, MICRO)

    def saStep(self, a, b, U, nnt, parity, temperature, rngSeed, rngCounter):
        self._sweep(a, b, U, None, parity, SA, temperature, rngSeed, rngCounter)

    def randomTrafo(self, a, b, U, nnt, parity, rngSeed, rngCounter):
        self._sweep(a, b, U, None, parity, RANDOM, 0.0, rngSeed, rngCounter)

    def generateGaugeQualityPerSite(self, a, b, U, dGff, dA):
        if a * b != self.lattice.V:
            raise GFError(_lib.GF_ERR_BAD_SHAPE, f"launch {a}x{b} does not cover V = {self.lattice.V} sites")
        _lib.call("gf_quality_per_site", C.byref(self.lattice.struct), self.gauge, _ptr(U), _ptr(dGff), _ptr(dA),
                  _ptr(self.stream))

    @classmethod
    def getGaugeQualityPrefactorA(cls) -> float:
        return 1.0 / cls.Nc

    @classmethod
    def getGaugeQualityPrefactorGff(cls) -> float:
        return 1.0 / (cls.Nc * cls.Ndim)


class LandauKernelsSU3(_GaugeKernels):
    gauge = LANDAU


class U1xU1KernelsSU3(_GaugeKernels):
    gauge = U1XU1


class MAGKernelsSU3(_GaugeKernels):
    gauge = MAG

    def srStep(self, a, b, U, nnt, parity, srParameter, rngSeed, rngCounter):
        self._sweep(a, b, U, None, parity, SR, srParameter, rngSeed, rngCounter)


class CoulombKernelsSU3(_GaugeKernels):
    """Coulomb sweeps of one timeslice: ``UtUp`` = slice t, ``UtDw`` = slice
    t-1 (CoulombKernelsSU3.hxx:208-250); the lattice is the timeslice
    (dims (1, Nx, Ny, Nz))."""

    gauge = COULOMB

    @classmethod
    def getGaugeQualityPrefactorGff(cls) -> float:
        return 1.0 / (cls.Nc * (cls.Ndim - 1))

    def orStep(self, a, b, UtUp, UtDw, nnt, parity, orParameter):
        self._sweep(a, b, UtUp, UtDw, parity, OR, orParameter)

    def microStep(self, a, b, UtUp, UtDw, nnt, parity):
        self._sweep(a, b, UtUp, UtDw, parity, MICRO)

    def saStep(self, a, b, UtUp, UtDw, nnt, parity, temperature, rngSeed, rngCounter):
        self._sweep(a, b, UtUp, UtDw, parity, SA, temperature, rngSeed, rngCounter)

    def randomTrafo(self, a, b, UtUp, UtDw, nnt, parity, rngSeed, rngCounter):
        self._sweep(a, b, UtUp, UtDw, parity, RANDOM, 0.0, rngSeed, rngCounter)


class CommonKernelsSU3:
    """projectSU3 / setHot (CommonKernelsSU3.hxx:40-62) on ``lattice``."""

    def __init__(self, lattice: Lattice, stream=None):
        self.lattice = lattice
        self.stream = stream

    def initCacheConfig(self) -> None:
        """No-op on gfx950."""

    def projectSU3(self, a, b, U, ptrToDeviceSize=None):
        if a * b != self.lattice.V:
            raise GFError(_lib.GF_ERR_BAD_SHAPE, f"launch {a}x{b} does not cover V = {self.lattice.V} sites")
        _lib.call("gf_project", C.byref(self.lattice.struct), _ptr(U), _ptr(self.stream))

    def setHot(self, a, b, U, ptrToDeviceSize, rngSeed, rngCounter):
        if a * b != self.lattice.V:
            raise GFError(_lib.GF_ERR_BAD_SHAPE, f"launch {a}x{b} does not cover V = {self.lattice.V} sites")
        _lib.call("gf_set_hot", C.byref(self.lattice.struct), _ptr(U), int(rngSeed),
                  C.c_uint32(rngCounter & 0xFFFFFFFF), _ptr(self.stream))


class GaugeFixingStats:
    """GaugeFixingStats<Ndim, Nc, GType, StoppingCrit> (GaugeFixingStats.hxx:55-232)."""

    def __init__(self, GType, U, lattice: Lattice, crit: int = CRIT_AVERAGE, prec: float | None = None, stream=None):
        self.gauge = GType.gauge if isinstance(GType, type) else int(GType)
        self.lattice = lattice
        self.U = U
        self.stream = stream
        self.reqPrec = prec
        self.currentGff = 0.0
        self.currentA = 0.0
        h = C.c_void_p()
        _lib.call("gf_stats_create", (C.c_int * 4)(*lattice.dims), self.gauge, crit, C.byref(h))
        self._h = h

    def generateGaugeQuality(self) -> None:
        g, a = C.c_double(), C.c_double()
        _lib.call("gf_stats_quality", self._h, C.byref(self.lattice.struct), _ptr(self.U), C.byref(g), C.byref(a),
                  _ptr(self.stream))
        self.currentGff, self.currentA = g.value, a.value

    def getCurrentGff(self) -> float:
        return self.currentGff

    def getCurrentA(self) -> float:
        return self.currentA

    def setPointer(self, U) -> None:
        self.U = U

    def setGaugePrecision(self, prec: float) -> None:
        self.reqPrec = prec

    def getGaugePrecision(self):
        return self.reqPrec

    def close(self) -> None:
        if self._h:
            _lib.load().gf_stats_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- drivers
@dataclass
class RunResult:
    sweeps: int
    sweeps_fp32: int
    checks: int
    gff: float
    theta: float
    ms_total: float
    ms_fp32: float
    converged: int


def _result(r: _lib.gf_run_result) -> RunResult:
    return RunResult(r.sweeps, r.sweeps_fp32, r.checks, r.gff, r.theta, r.ms_total, r.ms_fp32, int(r.converged))


def run_or(lattice: Lattice, U, gauge: int = LANDAU, omega: float = 1.7, precision: float = 1e-7,
           maxiter: int = 1000, checkprecision: int = 100, reproject: int = 100, stream=None) -> RunResult:
    """OR loop with the reference app's control flow (LandauGaugeFixingSU3_4D.cu:244-262)."""
    r = _lib.gf_run_result()
    _lib.call("gf_run_or", C.byref(lattice.struct), _ptr(U), gauge, C.c_float(omega), C.c_float(precision),
              int(maxiter), int(checkprecision), int(reproject), C.byref(r), _ptr(stream))
    return _result(r)


def run_fast(lattice: Lattice, U, gauge: int = LANDAU, omega: float = 1.7, precision: float = 1e-12,
             maxiter: int = 20000, reproject: int = 100, stream=None) -> RunResult:
    """OR with the check fused into the sweep kernels (every sweep checked,
    no host synchronisation per sweep); see gf_run_fast in include/su3gf.h."""
    r = _lib.gf_run_result()
    _lib.call("gf_run_fast", C.byref(lattice.struct), _ptr(U), gauge, C.c_float(omega), C.c_float(precision),
              int(maxiter), int(reproject), C.byref(r), _ptr(stream))
    return _result(r)


def run_fast_noalloc(lattice: Lattice, U, gauge: int = LANDAU, omega: float = 1.7, precision: float = 1e-12,
                     maxiter: int = 20000, reproject: int = 100, stream=None) -> RunResult:
    """run_fast without any allocation (its state lives in a fixed device
    array of the library: at most 64^4 sites; calls overlapping on one GPU
    race); see gf_run_fast_noalloc in include/su3gf.h."""
    r = _lib.gf_run_result()
    _lib.call("gf_run_fast_noalloc", C.byref(lattice.struct), _ptr(U), gauge, C.c_float(omega),
              C.c_float(precision), int(maxiter), int(reproject), C.byref(r), _ptr(stream))
    return _result(r)


def run_sa(lattice: Lattice, U, gauge: int = LANDAU, tmax: float = 0.4, tmin: float = 0.01, steps: int = 1000,
           micro: int = 3, reproject: int = 100, seed: int = 1, counter: int = 0, Udw=None, mode: int = 0,
           stream=None, graph: bool = False) -> tuple[int, int]:
    """Simulated annealing with the reference apps' schedule for ``gauge``
    (gf_run_sa in include/su3gf.h); mode 0 sweeps U itself, mode 1 anneals
    the gauge transformation in FP32 and applies it to U, mode 2 does the
    same on a copy of U in 24-bit fixed point; ``graph``: the schedule's
    launches are captured into one HIP graph and replayed.  Returns (next RNG
    counter, number of temperatures visited)."""
    ctr = C.c_uint32(counter)
    ntemps = C.c_int()
    _lib.call("gf_run_sa", C.byref(lattice.struct), _ptr(U), _ptr(Udw), gauge, C.c_float(tmax), C.c_float(tmin),
              int(steps), int(micro), int(reproject), int(mode) | (SA_GRAPH if graph else 0), int(seed),
              C.byref(ctr), C.byref(ntemps), _ptr(stream))
    return int(ctr.value), int(ntemps.value)


def _batch_args(lattices, Us, streams, max_concurrent):
    if len(lattices) != len(Us):
        raise ValueError("one configuration per lattice")
    n = len(lattices)
    lats = (_lib.gf_lattice * n)(*[l.struct for l in lattices])
    ptrs = (C.c_void_p * n)(*[_ptr(u).value for u in Us])
    strs = None
    if streams is not None:
        if len(streams) < max(1, max_concurrent):
            raise ValueError("one stream per concurrent problem")
        strs = (C.c_void_p * len(streams))(*[_ptr(s).value for s in streams])
    return n, lats, ptrs, strs


def batch_run_sa(lattices, Us, gauge: int = LANDAU, tmax: float = 0.4, tmin: float = 0.01, steps: int = 1000,
                 micro: int = 3, reproject: int = 100, seeds=None, counters=None, mode: int = 0,
                 max_concurrent: int = 4, streams=None, graph: bool = False,
                 one_launch: bool = False) -> list[tuple[int, int]]:
    """run_sa on independent 4D lattices in one call (gf_batch_run_sa in
    include/su3gf.h): small lattices concurrently, each on its own stream
    (``streams``: the caller's, one per concurrent problem, or None).
    ``seeds`` and ``counters`` give each problem's RNG seed and starting
    launch counter.  ``one_launch`` (modes 1 and 2): each half-sweep of a
    group of lattices is one launch (GF_SA_ONE_LAUNCH; groups limited by
    set_batch_launch_blocks).  Returns each problem's (next RNG counter,
    number of temperatures)."""
    n, lats, ptrs, strs = _batch_args(lattices, Us, streams, max_concurrent)
    seeds = list(seeds) if seeds is not None else [1] * n
    counters = list(counters) if counters is not None else [0] * n
    if len(seeds) != n or len(counters) != n:
        raise ValueError("one seed and one counter per problem")
    cs = (C.c_int * n)(*seeds)
    ctr = (C.c_uint32 * n)(*counters)
    temps = (C.c_int * n)()
    _lib.call("gf_batch_run_sa", n, lats, ptrs, gauge, C.c_float(tmax), C.c_float(tmin), int(steps), int(micro),
              int(reproject), int(mode) | (SA_GRAPH if graph else 0) | (SA_ONE_LAUNCH if one_launch else 0), cs, ctr,
              temps, int(max_concurrent), strs)
    return [(int(ctr[k]), int(temps[k])) for k in range(n)]


def set_batch_launch_blocks(blocks: int) -> int:
    """Limit of the 64-site blocks of one parity that a one-launch batch
    annealing launch gathers (gf_set_batch_launch_blocks; 0: no limit but 16
    lattices).  Returns the previous limit."""
    prev = C.c_int()
    _lib.call("gf_set_batch_launch_blocks", int(blocks), C.byref(prev))
    return prev.value


def batch_run_mixed(lattices, Us, gauge: int = LANDAU, omega: float = 1.7, switch_precision: float = 1e-6,
                    precision: float = 1e-12, maxiter: int = 20000, mode: int | None = None, refresh: int = 50,
                    check32: int = 10, reproject: int = 100, max_concurrent: int = 4,
                    streams=None) -> list[RunResult]:
    """run_mixed on independent 4D lattices in one call (gf_batch_run_mixed in
    include/su3gf.h): small lattices concurrently, each on its own stream
    (``streams``: the caller's, one per concurrent problem, or None)."""
    if mode is None:
        mode = MIXED_GAUGE24
    n, lats, ptrs, strs = _batch_args(lattices, Us, streams, max_concurrent)
    out = (_lib.gf_run_result * n)()
    _lib.call("gf_batch_run_mixed", n, lats, ptrs, gauge, C.c_float(omega), C.c_float(switch_precision),
              C.c_float(precision), int(maxiter), int(mode), int(refresh), int(check32), int(reproject),
              int(max_concurrent), strs, out)
    return [_result(out[k]) for k in range(n)]


def run_mixed(lattice: Lattice, U, gauge: int = LANDAU, omega: float = 1.7, switch_precision: float = 1e-6,
              precision: float = 1e-12, maxiter: int = 20000, mode: int | None = None, refresh: int = 50,
              check32: int = 10, reproject: int = 100, stream=None) -> RunResult:
    """FP32 sweeps accumulating a gauge transformation down to
    ``switch_precision`` (checked every ``check32`` sweeps), the transformation
    applied to the FP64 configuration, then FP64 fused-check sweeps to
    ``precision``.  ``mode``: MIXED_GAUGE24 (default), MIXED_GAUGE32,
    MIXED_GAUGE64, MIXED_LINKS, or MIXED_FP64 (no FP32 phase: the FP64 phase
    alone) -- see gf_run_mixed in include/su3gf.h."""
    if mode is None:
        mode = MIXED_GAUGE24
    r = _lib.gf_run_result()
    _lib.call("gf_run_mixed", C.byref(lattice.struct), _ptr(U), gauge, C.c_float(omega), C.c_float(switch_precision),
              C.c_float(precision), int(maxiter), int(mode), int(refresh), int(check32), int(reproject), C.byref(r),
              _ptr(stream))
    return _result(r)


# ------------------------------------------------------------ files
class LinkFile:
    """A configuration file of one type (include/su3gf_io.h), as the
    reference's LinkFile<FileType, StandardPattern, GpuPattern, SiteCoord>
    (lattice/LinkFile.hxx:60-72) and the apps' ILDG / QCDSTAG functions.
    ``U`` is a host numpy array of all 18 reals per link (float64 or float32)
    in GpuPattern order (or GpuPatternTimeslice with
    ``pattern=PATTERN_GPU_TIMESLICE``).  The handle keeps the header or LIME
    records of the last file loaded, for ``save``."""

    def __init__(self, filetype: int, reinterpret: int = REAL_STANDARD, pattern: int = PATTERN_GPU):
        h = C.c_void_p()
        _lib.call("gf_linkfile_create", int(filetype), int(reinterpret), int(pattern), C.byref(h))
        self._h = h
        self.filetype = filetype

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().gf_linkfile_destroy(h)
            self._h = None

    @staticmethod
    def _lattice(dims, U: np.ndarray) -> "Lattice":
        dtype = {np.dtype(np.float64): F64, np.dtype(np.float32): F32}.get(U.dtype)
        if dtype is None:
            raise TypeError("U must be float64 or float32")
        return Lattice(tuple(dims), dtype, PATTERN18)

    def load(self, dims, filename: str, U: np.ndarray) -> None:
        """Fill ``U`` (contiguous, 72 V reals) from ``filename``."""
        if not (U.flags.c_contiguous and U.flags.writeable):
            raise ValueError("U must be a writeable contiguous array")
        lat = self._lattice(dims, U)
        if U.size != 72 * lat.V:
            raise ValueError(f"U holds {U.size} reals, the lattice {72 * lat.V}")
        _lib.call("gf_linkfile_load", self._h, str(filename).encode(), C.byref(lat.struct), U.ctypes.data_as(C.c_void_p))

    def save(self, dims, filename: str, U: np.ndarray, sa_steps: int = 0) -> None:
        U = np.ascontiguousarray(U)
        lat = self._lattice(dims, U)
        if U.size != 72 * lat.V:
            raise ValueError(f"U holds {U.size} reals, the lattice {72 * lat.V}")
        _lib.call("gf_linkfile_save", self._h, str(filename).encode(), C.byref(lat.struct),
                  U.ctypes.data_as(C.c_void_p), int(sa_steps))

    def read_header(self, dims, filename: str, dtype=np.float64) -> None:
        """Keep the header / LIME records of ``filename`` without its links."""
        lat = Lattice(tuple(dims), F64 if np.dtype(dtype) == np.float64 else F32, PATTERN18)
        _lib.call("gf_linkfile_read_header", self._h, str(filename).encode(), C.byref(lat.struct))


# --------------------------------------------------------- synthetic input
def make_config(dims, sigma: float, seed: int, haar_seed: int = 0) -> np.ndarray:
    """Host FP64 configuration in the 18-real GpuPattern layout (V*72 reals):
    U = exp(i sum_a c_a lambda_a/2), c_a ~ N(0, sigma^2), then a Haar gauge
    transformation when ``haar_seed`` != 0."""
    dims = tuple(int(d) for d in dims)
    V = dims[0] * dims[1] * dims[2] * dims[3]
    out = np.empty(V * 72, dtype=np.float64)
    _lib.check(_lib.load().gf_make_config((C.c_int * 4)(*dims), float(sigma), C.c_uint32(seed),
                                          C.c_uint32(haar_seed), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out
