"""The gauge-fixing applications' schedule, run on the CPU oracle or on the
gfx950 library through its Python mirror.

Test infrastructure: this restates, step for step, the per-copy loops of the
reference applications -- LandauGaugeFixingSU3_4D.cu:184-286,
U1xU1GaugeFixingSU3_4D.cu:190-300, MAGaugeFixingSU3_4D.cu:199-345 and
CoulombGaugeFixingSU3_4D.cu:198-310 -- so that the application built on the
drop-in headers (su3_ma_gauge_gpu_amd/apps) can be compared with it RESULT
line by RESULT line.  The RNG launch counter, the float temperature
arithmetic and the float/double comparisons follow the reference exactly,
since they decide how many sweeps run.

Two backends run the same schedule: OracleBackend (the CPU restatement of the
sweep) and GpuBackend (the library's gf_sweep / gf_project / gf_stats calls,
the kernels the application runs).  Heatbath schedules are chaotic at the
level of the last bit of log / cos / sin, so they are compared with the
application exactly on the GpuBackend and with the oracle only where the
dynamics contracts.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle_lib import COULOMB, LANDAU, MAG, MICRO, OR, RANDOM, SA, SR, U1XU1

GAUGES = {"landau": LANDAU, "u1xu1": U1XU1, "mag": MAG, "coulomb": COULOMB}


@dataclass
class Options:
    """apps/program_options/ProgramOptions.hxx:229-266 defaults."""
    gaugecopies: int = 1
    randomtrafo: int = 1
    reproject: int = 100
    seed: int = 1
    sasteps: int = 1000
    samin: float = 0.01
    samax: float = 0.4
    microupdates: int = 3
    ormaxiter: int = 1000
    orparameter: float = 1.7
    srmaxiter: int = 1000
    srparameter: float = 1.7
    precision: float = 1e-7
    checkprecision: int = 100
    doSA: int = 1
    hotgaugefield: int = 0

    def argv(self) -> list[str]:
        out = []
        for k, v in self.__dict__.items():
            out += [f"--{k}", repr(v) if isinstance(v, float) else str(v)]
        return out


class Counter:
    """PhiloxWrapper::getNextCounter (lattice/rng/PhiloxWrapper.hxx)."""

    def __init__(self):
        self.n = 0

    def next(self) -> int:
        self.n += 1
        return self.n - 1


F = np.float32


class OracleBackend:
    """Schedule steps on the CPU oracle; arrays are numpy link arrays."""

    def __init__(self, oracle):
        self.o = oracle

    def array(self, host):
        return np.array(host, dtype=self.o.np_real)

    def empty(self, n):
        return np.zeros(n, dtype=self.o.np_real)

    def host(self, a):
        return a

    def sweep(self, g, algo, dims, U, parity, param=0.0, seed=0, counter=0):
        self.o.sweep(g, algo, dims, U, parity, param, seed, counter)

    def sweep_coulomb(self, algo, dims3, up, dw, parity, param=0.0, seed=0, counter=0):
        self.o.sweep_coulomb(algo, dims3, up, dw, parity, param, seed, counter)

    def project_all(self, dims, U):
        self.o.project_all(dims, U)

    def set_hot(self, dims, U, seed, counter):
        self.o.set_hot(dims, U, seed, counter)

    def quality(self, g, dims, U):
        return self.o.quality(g, dims, U)


class GpuBackend:
    """Schedule steps through the library (the calls the application's
    headers make); arrays are device buffers of FP64 GpuPattern links."""

    class Arr:
        def __init__(self, buf, n):
            self.buf, self.n = buf, n

    def __init__(self, gf):
        self.gf = gf
        self._stats = {}

    def _lat(self, dims):
        return self.gf.Lattice(dims)

    def array(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        return self.Arr(self.gf.DeviceBuffer.from_numpy(host), host.size)

    def empty(self, n):
        return self.array(np.zeros(n))

    def host(self, a):
        return a.buf.download(np.float64, a.n)

    def sweep(self, g, algo, dims, U, parity, param=0.0, seed=0, counter=0):
        import ctypes as C
        lat = self._lat(dims)
        self.gf._lib.call("gf_sweep", C.byref(lat.struct), C.c_void_p(U.buf.ptr), C.c_void_p(U.buf.ptr), parity, g,
                          algo, C.c_float(param), seed, C.c_uint32(counter & 0xFFFFFFFF), 32, None)

    def sweep_coulomb(self, algo, dims3, up, dw, parity, param=0.0, seed=0, counter=0):
        import ctypes as C
        lat = self._lat(dims3)
        self.gf._lib.call("gf_sweep", C.byref(lat.struct), C.c_void_p(up.buf.ptr), C.c_void_p(dw.buf.ptr), parity,
                          COULOMB, algo, C.c_float(param), seed, C.c_uint32(counter & 0xFFFFFFFF), 32, None)

    def project_all(self, dims, U):
        self.gf.CommonKernelsSU3(self._lat(dims)).projectSU3(int(np.prod(dims)) // 32, 32, U.buf)

    def set_hot(self, dims, U, seed, counter):
        self.gf.CommonKernelsSU3(self._lat(dims)).setHot(int(np.prod(dims)) // 32, 32, U.buf, None, seed, counter)

    def quality(self, g, dims, U):
        key = (g, tuple(dims))
        if key not in self._stats:
            self._stats[key] = self.gf.GaugeFixingStats(g, None, self._lat(dims))
        st = self._stats[key]
        st.setPointer(U.buf)
        st.generateGaugeQuality()
        return st.getCurrentGff(), st.getCurrentA()


def as_backend(x):
    return x if isinstance(x, (OracleBackend, GpuBackend)) else OracleBackend(x)


def run_4d(backend, gauge: str, dims, U0: np.ndarray, o: Options) -> list[dict]:
    oracle = as_backend(backend)
    g = GAUGES[gauge]
    mag, u1 = g == MAG, g == U1XU1
    prec = float(F(o.precision))
    ctr = Counter()
    results = []
    U = oracle.array(U0)
    host_U, best = U0, 0.0  # the app's host array: the best copy so far is copied back into it
    if o.hotgaugefield:  # once, before the copies; the copies then continue from each other
        oracle.set_hot(dims, U, o.seed, ctr.next())
    for copy in range(o.gaugecopies):
        if not o.hotgaugefield:
            U = oracle.array(host_U)
        if o.randomtrafo:
            for parity in (0, 1):
                oracle.sweep(g, RANDOM, dims, U, parity, 0.0, o.seed, ctr.next())
        gff, theta = oracle.quality(g, dims, U)

        def sa_sweep(temperature):
            for parity in (0, 1):
                oracle.sweep(g, SA, dims, U, parity, float(temperature), o.seed, ctr.next())
            for _ in range(o.microupdates):
                for parity in (0, 1):
                    oracle.sweep(g, MICRO, dims, U, parity)

        if not mag or o.doSA:
            temperature = F(o.samax)
            base = F(F(o.samax) - F(o.samin)) / F(o.sasteps) if o.sasteps else F(np.inf)
            if not mag and not u1:
                for i in range(o.sasteps):
                    sa_sweep(temperature)
                    if i % o.reproject == 0:
                        oracle.project_all(dims, U)
                    if i % o.checkprecision == 0:
                        gff, theta = oracle.quality(g, dims, U)
                    temperature = F(temperature - base)
            else:
                lo, hi = (2.0, 2.7) if u1 else (0.7, 0.88)
                i = 0
                while True:
                    step = F(base / F(15)) if (float(temperature) <= hi and float(temperature) >= lo) else base
                    for _ in range(5 if mag else 1):
                        sa_sweep(temperature)
                    if i % o.reproject == 0:
                        oracle.project_all(dims, U)
                    if u1 and i % o.checkprecision == 0:
                        gff, theta = oracle.quality(g, dims, U)
                    temperature = F(temperature - step)
                    i += 1
                    if not temperature >= F(o.samin):
                        break

        sr_steps = 0
        if mag:
            for i in range(o.srmaxiter):
                for parity in (0, 1):
                    oracle.sweep(g, SR, dims, U, parity, o.srparameter, o.seed, ctr.next())
                sr_steps = i + 1
                if i % o.reproject == 0:
                    oracle.project_all(dims, U)
                if i % o.checkprecision == 0:
                    gff, theta = oracle.quality(g, dims, U)
                    if theta < prec or theta < 1e-6:
                        break

        or_steps = 0
        for i in range(o.ormaxiter):
            for parity in (0, 1):
                oracle.sweep(g, OR, dims, U, parity, o.orparameter)
            or_steps = i + 1
            if i % o.reproject == 0:
                oracle.project_all(dims, U)
            if i % o.checkprecision == 0:
                gff, theta = oracle.quality(g, dims, U)
                if theta < prec:
                    break
        oracle.project_all(dims, U)
        if mag:
            gff, theta = oracle.quality(g, dims, U)
        results.append(dict(copy=copy, gff=gff, theta=theta, or_sweeps=or_steps, sr_sweeps=sr_steps,
                            counter=ctr.n, U=oracle.host(U)))
        if gff > best:  # LandauGaugeFixingSU3_4D.cu:274-283: copy back the better copy
            best = gff
            host_U = results[-1]["U"]
    return results


def to_timeslices(U4: np.ndarray, dims) -> np.ndarray:
    """GpuPattern (4D, parity split over the whole lattice) -> one
    GpuPatternTimeslice array per t (parity split within the slice)."""
    Nt, Nx, Ny, Nz = dims
    V, V3 = Nt * Nx * Ny * Nz, Nx * Ny * Nz
    t, x, y, z = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    n4 = ((t * Nx + x) * Ny + y) * Nz + z
    i4 = np.where((t + x + y + z) % 2 == 0, n4 // 2, n4 // 2 + V // 2)
    n3 = (x * Ny + y) * Nz + z
    i3 = np.where((x + y + z) % 2 == 0, n3 // 2, n3 // 2 + V3 // 2)
    A = U4.reshape(72, V)
    out = np.empty((Nt, 72, V3), dtype=U4.dtype)
    out[t.ravel(), :, i3.ravel()] = A[:, i4.ravel()].T
    return out


def run_coulomb(backend, dims, U4: np.ndarray, o: Options) -> list[dict]:
    oracle = as_backend(backend)
    Nt = dims[0]
    dims3 = (1, dims[1], dims[2], dims[3])
    slices = to_timeslices(U4, dims)
    prec = float(F(o.precision))
    ctr = Counter()
    results = []
    for t in range(Nt):
        tdw = t - 1 if t > 0 else Nt - 1
        best = 0.0
        for copy in range(o.gaugecopies):
            if not o.hotgaugefield:
                up = oracle.array(slices[t].reshape(-1))
                dw = oracle.array(slices[tdw].reshape(-1))
            else:
                up = oracle.empty(slices[t].size)
                dw = oracle.empty(slices[t].size)
                oracle.set_hot(dims3, up, o.seed, ctr.next())
                oracle.set_hot(dims3, dw, o.seed, ctr.next())
            if o.randomtrafo:
                for parity in (0, 1):
                    oracle.sweep_coulomb(RANDOM, dims3, up, dw, parity, 0.0, o.seed, ctr.next())
            gff, theta = oracle.quality(COULOMB, dims3, up)
            temperature = F(o.samax)
            step = F(F(o.samax) - F(o.samin)) / F(o.sasteps) if o.sasteps else F(np.inf)
            for i in range(o.sasteps):
                for parity in (0, 1):
                    oracle.sweep_coulomb(SA, dims3, up, dw, parity, float(temperature), o.seed, ctr.next())
                for _ in range(o.microupdates):
                    for parity in (0, 1):
                        oracle.sweep_coulomb(MICRO, dims3, up, dw, parity)
                if i % o.reproject == 0:
                    oracle.project_all(dims3, up)
                    oracle.project_all(dims3, dw)
                if i % o.checkprecision == 0:
                    gff, theta = oracle.quality(COULOMB, dims3, up)
                temperature = F(temperature - step)
            or_steps = 0
            for i in range(o.ormaxiter):
                for parity in (0, 1):
                    oracle.sweep_coulomb(OR, dims3, up, dw, parity, o.orparameter)
                or_steps = i + 1
                if i % o.reproject == 0:
                    oracle.project_all(dims3, up)
                    oracle.project_all(dims3, dw)
                if i % o.checkprecision == 0:
                    gff, theta = oracle.quality(COULOMB, dims3, up)
                    if theta < prec:
                        break
            oracle.project_all(dims3, up)
            oracle.project_all(dims3, dw)
            results.append(dict(slice=t, copy=copy, gff=gff, theta=theta, or_sweeps=or_steps, counter=ctr.n))
            if gff > best:
                best = gff
                slices[t] = oracle.host(up).reshape(72, -1)
                slices[tdw] = oracle.host(dw).reshape(72, -1)
    return results


def parse_results(stdout: str) -> list[dict]:
    out = []
    for line in stdout.splitlines():
        if line.startswith("RESULT "):
            d = {}
            for kv in line.split()[1:]:
                k, v = kv.split("=")
                d[k] = float(v) if k in ("gff", "theta") else int(v)
            out.append(d)
    return out
