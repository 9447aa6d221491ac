"""ctypes binding of ``libsu3gf.so`` (the C ABI declared in include/su3gf.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or does not load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsu3gf.so"

GF_OK, GF_ERR_BAD_SHAPE, GF_ERR_BAD_ARG, GF_ERR_HIP, GF_ERR_NOT_CONVERGED, GF_ERR_NO_DEVICE, GF_ERR_IO = range(7)
LANDAU, COULOMB, MAG, U1XU1 = range(4)
OR, MICRO, SA, SR, RANDOM = range(5)
UPDATE_DP = 0x100  # OR'ed into an update rule: computed in double on FP32 links
SA_GRAPH = 0x10  # OR'ed into an annealing mode: the schedule captured into one HIP graph
SA_ONE_LAUNCH = 0x20  # OR'ed into a batch annealing mode: each half-sweep of several lattices in one launch
F32, F64 = 0, 1
R12, PATTERN18 = 12, 18
MIXED_LINKS, MIXED_GAUGE32, MIXED_GAUGE64, MIXED_GAUGE24, MIXED_FP64 = 0, 1, 2, 3, 4
WAVEFRONT_PIPELINED = 1000000  # gf_set_wavefront: the pipelined sweep, one launch per sweep
CRIT_MAX, CRIT_AVERAGE = 0, 1
FILE_PLAIN, FILE_HEADERONLY, FILE_VOGT, FILE_ILDG, FILE_QCDSTAG = range(5)
REAL_STANDARD, REAL_DOUBLE, REAL_FLOAT = range(3)
PATTERN_GPU, PATTERN_GPU_TIMESLICE = 0, 1

_STATUS_NAMES = {
    GF_ERR_BAD_SHAPE: "BAD_SHAPE",
    GF_ERR_BAD_ARG: "BAD_ARG",
    GF_ERR_HIP: "HIP_ERROR",
    GF_ERR_NOT_CONVERGED: "NOT_CONVERGED",
    GF_ERR_NO_DEVICE: "NO_DEVICE",
    GF_ERR_IO: "IO_ERROR",
}


class GFError(RuntimeError):
    """A gf_status other than GF_OK, with the library's error text."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class gf_lattice(C.Structure):
    _fields_ = [("dims", C.c_int * 4), ("dtype", C.c_int), ("layout", C.c_int)]


class gf_run_result(C.Structure):
    _fields_ = [
        ("sweeps", C.c_int),
        ("sweeps_fp32", C.c_int),
        ("checks", C.c_int),
        ("gff", C.c_double),
        ("theta", C.c_double),
        ("ms_total", C.c_double),
        ("ms_fp32", C.c_double),
        ("converged", C.c_int),
    ]


_vp = C.c_void_p
_i = C.c_int
_u32 = C.c_uint32
_f = C.c_float
_d = C.c_double
_pl = C.POINTER(gf_lattice)
_pi4 = C.POINTER(C.c_int)

# name -> (restype, argtypes); every function declared in include/*.h
SIGNATURES = {
    "gf_last_error": (C.c_char_p, []),
    "gf_version": (C.c_char_p, []),
    "gf_device_info": (_i, [C.POINTER(_i), C.c_char_p, C.c_size_t]),
    "gf_malloc": (_i, [C.POINTER(_vp), C.c_size_t]),
    "gf_free": (_i, [_vp]),
    "gf_memcpy_h2d": (_i, [_vp, _vp, C.c_size_t]),
    "gf_memcpy_d2h": (_i, [_vp, _vp, C.c_size_t]),
    "gf_memcpy_d2d": (_i, [_vp, _vp, C.c_size_t]),
    "gf_memcpy_async": (_i, [_vp, _vp, C.c_size_t, _vp]),
    "gf_memset": (_i, [_vp, _i, C.c_size_t]),
    "gf_synchronize": (_i, []),
    "gf_stream_create": (_i, [C.POINTER(_vp)]),
    "gf_stream_destroy": (_i, [_vp]),
    "gf_stream_synchronize": (_i, [_vp]),
    "gf_event_create": (_i, [C.POINTER(_vp)]),
    "gf_event_destroy": (_i, [_vp]),
    "gf_event_record": (_i, [_vp, _vp]),
    "gf_event_elapsed_ms": (_i, [_vp, _vp, C.POINTER(_f)]),
    "gf_lattice_bytes": (C.c_size_t, [_pl]),
    "gf_sweep": (_i, [_pl, _vp, _vp, _i, _i, _i, _f, _i, _u32, _i, _vp]),
    "gf_project": (_i, [_pl, _vp, _vp]),
    "gf_set_hot": (_i, [_pl, _vp, _i, _u32, _vp]),
    "gf_convert": (_i, [_pl, _vp, _pl, _vp, _i, _vp]),
    "gf_links24_words": (C.c_size_t, [_pi4]),
    "gf_pack_links24": (_i, [_pl, _vp, _vp, _vp]),
    "gf_unpack_links24": (_i, [_pi4, _vp, _vp, _vp]),
    "gf_stats_create": (_i, [_pi4, _i, _i, C.POINTER(_vp)]),
    "gf_stats_destroy": (_i, [_vp]),
    "gf_stats_quality": (_i, [_vp, _pl, _vp, C.POINTER(_d), C.POINTER(_d), _vp]),
    "gf_quality_per_site": (_i, [_pl, _i, _vp, _vp, _vp, _vp]),
    "gf_run_or": (_i, [_pl, _vp, _i, _f, _f, _i, _i, _i, C.POINTER(gf_run_result), _vp]),
    "gf_run_fast": (_i, [_pl, _vp, _i, _f, _f, _i, _i, C.POINTER(gf_run_result), _vp]),
    "gf_run_fast_noalloc": (_i, [_pl, _vp, _i, _f, _f, _i, _i, C.POINTER(gf_run_result), _vp]),
    "gf_time_sweeps": (_i, [_pl, _vp, _i, _i, _i, C.POINTER(_f), _vp]),
    "gf_set_wavefront": (_i, [_i, C.POINTER(_i)]),
    "gf_wavefront_choice": (_i, [_pl, _i, _i, C.POINTER(_i)]),
    "gf_pipelined_launches": (_i, [C.POINTER(C.c_longlong)]),
    "gf_set_pipelined_lag": (_i, [_i, C.POINTER(_i)]),
    "gf_set_block_order": (_i, [_i, C.POINTER(_i)]),
    "gf_set_batch_launch_blocks": (_i, [_i, C.POINTER(_i)]),
    "gf_block_order_choice": (_i, [_pl, _i, _i, C.POINTER(_i)]),
    "gf_run_sa": (_i, [_pl, _vp, _vp, _i, _f, _f, _i, _i, _i, _i, _i, C.POINTER(C.c_uint32), C.POINTER(_i), _vp]),
    "gf_run_mixed": (_i, [_pl, _vp, _i, _f, _f, _f, _i, _i, _i, _i, _i, C.POINTER(gf_run_result), _vp]),
    "gf_batch_run_mixed": (_i, [_i, _pl, C.POINTER(_vp), _i, _f, _f, _f, _i, _i, _i, _i, _i, _i, C.POINTER(_vp),
                                C.POINTER(gf_run_result)]),
    "gf_batch_run_sa": (_i, [_i, _pl, C.POINTER(_vp), _i, _f, _f, _i, _i, _i, _i, C.POINTER(_i),
                             C.POINTER(C.c_uint32), C.POINTER(_i), _i, C.POINTER(_vp)]),
    "gf_make_config": (_i, [_pi4, _d, _u32, _u32, C.POINTER(_d)]),
    # include/su3gf_io.h
    "gf_linkfile_create": (_i, [_i, _i, _i, C.POINTER(_vp)]),
    "gf_linkfile_destroy": (None, [_vp]),
    "gf_linkfile_load": (_i, [_vp, C.c_char_p, _pl, _vp]),
    "gf_linkfile_save": (_i, [_vp, C.c_char_p, _pl, _vp, _i]),
    "gf_linkfile_read_header": (_i, [_vp, C.c_char_p, _pl]),
}

_lib = None


def load() -> C.CDLL:
    """Load libsu3gf.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m su3_ma_gauge_gpu_amd.build` "
                "(the HIP library is the only implementation)"
            )
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    msg = load().gf_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status != GF_OK:
        raise GFError(status, last_error())


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
