"""The 24-bit fixed-point link copy: the numpy restatement against hand-derived
words (CPU), and the device packing and decode against it bit for bit (GPU)."""
import ctypes as C

import numpy as np
import pytest

from links24 import Q, pack_config, pack_words, quantise, unpack_config, unpack_ints, words


def test_known_words():
    # 1 -> 0x7fffff, -1 -> 0x800001, 0.5 -> 4194303.5 rounds to even 0x400000,
    # -1e-7 -> -0.84 rounds to -1 = 0xffffff
    q = quantise([1.0, -1.0, 0.5, -1e-7])
    assert list(q) == [Q, -Q, 4194304, -1]
    w = pack_words(q)
    assert [int(x) for x in w] == [0x017FFFFF, 0x00008000, 0xFFFFFF40]
    assert list(unpack_ints(w)) == [Q, -Q, 4194304, -1]
    # clamping
    assert list(quantise([1.5, -3.0])) == [Q, -Q]


def test_round_trip_error():
    rng = np.random.default_rng(5)
    q = quantise(rng.uniform(-1, 1, size=(20000, 4)))
    assert np.array_equal(unpack_ints(pack_words(q)), q)
    v = rng.uniform(-1, 1, size=(4, 12 * 500))
    V = 500
    U = np.zeros((4, 18, V))
    U[:, :12, :] = v.reshape(4, 12, V)
    back = unpack_config(pack_config(U.reshape(-1), V), V).reshape(4, 12, V).astype(np.float64)
    # quantisation (2^-24) plus the FP32 product's rounding
    assert np.abs(back - U[:, :12, :]).max() <= 2.0 ** -24 + 2.0 ** -24


@pytest.mark.gpu
@pytest.mark.parametrize("cold,dims", [(False, (4, 6, 4, 8)), (True, (4, 6, 4, 8)), (False, (6, 6, 6, 6))])
def test_device_pack_and_decode_bit_exact(gf, gpu, cold, dims):
    V = int(np.prod(dims))
    U = gf.make_config(dims, 0.0 if cold else 0.4, 3, 0 if cold else 4)
    d = gf.DeviceBuffer.from_numpy(U)
    lat = gf.Lattice(dims)
    from su3_ma_gauge_gpu_amd import _lib
    n = _lib.load().gf_links24_words((C.c_int * 4)(*dims))
    assert n == words(V)
    P = gf.DeviceBuffer(4 * n)
    _lib.call("gf_memset", C.c_void_p(P.ptr), 0, C.c_size_t(4 * n))
    _lib.call("gf_pack_links24", C.byref(lat.struct), C.c_void_p(d.ptr), C.c_void_p(P.ptr), None)
    got_words = P.download(np.uint32)
    want = pack_config(U, V)
    assert np.array_equal(got_words, want)
    out = gf.DeviceBuffer(48 * V * 4)
    _lib.call("gf_unpack_links24", (C.c_int * 4)(*dims), C.c_void_p(P.ptr), C.c_void_p(out.ptr), None)
    got = out.download(np.float32)
    assert np.array_equal(got, unpack_config(want, V))
