"""ctypes binding of oracle/_ref/libref_io.so: the reference's own
configuration-file code (ildg.cpp, qcdstag.cpp, c-lime, FileVogt.hxx and the
pattern classes) compiled by oracle/Makefile.  Test infrastructure only."""
import ctypes as C
from pathlib import Path

import numpy as np

PATH = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libref_io.so"


class RefIO:
    def __init__(self):
        L = C.CDLL(str(PATH))
        P = C.POINTER(C.c_double)
        I4 = C.c_int * 4
        L.ref_read_ildg.argtypes = [I4, C.c_char_p, P, C.c_int]
        L.ref_write_ildg.argtypes = [I4, C.c_char_p, C.c_char_p, P, C.c_int, C.c_int]
        L.ref_read_qcdstag.argtypes = [I4, C.c_char_p, P, C.c_int]
        L.ref_write_qcdstag.argtypes = [I4, C.c_char_p, P, C.c_int]
        L.ref_index_by_unique.argtypes = [I4, C.c_longlong, C.c_int]
        L.ref_index_by_unique.restype = C.c_longlong
        L.ref_write_vogt.argtypes = [I4, C.POINTER(C.c_short), C.c_int, C.c_char_p, P]
        L.ref_read_vogt_header.argtypes = [C.c_char_p, C.POINTER(C.c_short)]
        L.ref_lime_write.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_char_p), C.POINTER(C.c_ulonglong)]
        L.ref_lime_list.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_ulonglong)]
        L.ref_lime_read.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_ulonglong]
        L.ref_lime_read.restype = C.c_longlong
        self.lib = L

    @staticmethod
    def _p(a):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def read_ildg(self, dims, path, V, timeslice=0):
        U = np.zeros(72 * V)
        self.lib.ref_read_ildg((C.c_int * 4)(*dims), str(path).encode(), self._p(U), timeslice)
        return U

    def write_ildg(self, dims, template, out, U, steps, timeslice=0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        self.lib.ref_write_ildg((C.c_int * 4)(*dims), str(template).encode(), str(out).encode(), self._p(U),
                                steps, timeslice)

    def read_qcdstag(self, dims, path, V, timeslice=0):
        U = np.zeros(72 * V)
        ok = self.lib.ref_read_qcdstag((C.c_int * 4)(*dims), str(path).encode(), self._p(U), timeslice)
        assert ok
        return U

    def write_qcdstag(self, dims, path, U, timeslice=0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        assert self.lib.ref_write_qcdstag((C.c_int * 4)(*dims), str(path).encode(), self._p(U), timeslice)

    def index_by_unique(self, dims, u, timeslice=0):
        return self.lib.ref_index_by_unique((C.c_int * 4)(*dims), int(u), timeslice)

    def write_vogt(self, dims, header, path, U):
        """header = (nc, latsize..., length_of_real)"""
        h = (C.c_short * len(header))(*header)
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert self.lib.ref_write_vogt((C.c_int * 4)(*dims), h, len(header) - 2, str(path).encode(), self._p(U))

    def read_vogt_header(self, path):
        out = (C.c_short * 24)()
        assert self.lib.ref_read_vogt_header(str(path).encode(), out)
        ndim = out[0]
        return [out[k] for k in range(3 + ndim)]

    def lime_write(self, path, records):
        """records: [(type, mb, me, bytes)]"""
        n = len(records)
        types = (C.c_char_p * n)(*[r[0].encode() for r in records])
        mb = (C.c_int * n)(*[int(r[1]) for r in records])
        me = (C.c_int * n)(*[int(r[2]) for r in records])
        data = (C.c_char_p * n)(*[bytes(r[3]) for r in records])
        nb = (C.c_ulonglong * n)(*[len(r[3]) for r in records])
        assert self.lib.ref_lime_write(str(path).encode(), n, types, mb, me, data, nb)

    def lime_list(self, path, max_records=32):
        types = C.create_string_buffer(128 * max_records)
        flags = (C.c_int * max_records)()
        nb = (C.c_ulonglong * max_records)()
        n = self.lib.ref_lime_list(str(path).encode(), max_records, types, flags, nb)
        assert n >= 0
        return [(types.raw[128 * k:128 * k + 128].split(b"\0")[0].decode(), bool(flags[k] & 2), bool(flags[k] & 1),
                 nb[k]) for k in range(n)]

    def lime_read(self, path, index, cap):
        buf = C.create_string_buffer(cap)
        n = self.lib.ref_lime_read(str(path).encode(), index, buf, cap)
        assert n >= 0
        return buf.raw[:n]
