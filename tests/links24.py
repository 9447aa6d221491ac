"""numpy restatement of the 24-bit fixed-point link copy (gf_pack_links24 /
gf_unpack_links24 in include/su3gf.h): the checker of the device packing."""
import numpy as np

Q = 8388607  # 2^23 - 1


def quantise(v):
    """round(v (2^23 - 1)) of v clamped to [-1, 1], ties to even, as int64."""
    return np.rint(np.clip(np.asarray(v, dtype=np.float64), -1.0, 1.0) * Q).astype(np.int64)


def pack_words(q):
    """Integers (..., 4) -> the three words (..., 3) of each group of four."""
    u = (np.asarray(q, dtype=np.int64) & 0xFFFFFF).astype(np.uint64)
    a, b, c, d = u[..., 0], u[..., 1], u[..., 2], u[..., 3]
    m = np.uint64(0xFFFFFFFF)
    w0 = (a | (b << np.uint64(24))) & m
    w1 = ((b >> np.uint64(8)) | (c << np.uint64(16))) & m
    w2 = ((c >> np.uint64(16)) | (d << np.uint64(8))) & m
    return np.stack([w0, w1, w2], axis=-1).astype(np.uint32)


def unpack_ints(w):
    """Words (..., 3) -> the signed integers (..., 4)."""
    w = np.asarray(w, dtype=np.uint32).astype(np.uint64)
    w0, w1, w2 = w[..., 0], w[..., 1], w[..., 2]
    a = w0 & np.uint64(0xFFFFFF)
    b = ((w0 >> np.uint64(24)) | (w1 << np.uint64(8))) & np.uint64(0xFFFFFF)
    c = ((w1 >> np.uint64(16)) | (w2 << np.uint64(16))) & np.uint64(0xFFFFFF)
    d = w2 >> np.uint64(8)
    u = np.stack([a, b, c, d], axis=-1).astype(np.int64)
    return np.where(u >= (1 << 23), u - (1 << 24), u)


def words(V):
    """Words of the copy: 9 per link in whole tiles of 64 sites."""
    return 4 * 9 * 64 * ((V + 63) // 64)


def pack_config(U, V):
    """FP64 links in the 18-real GpuPattern layout (V*72 reals) -> P: word w of
    link (i, mu) at P[576 T mu + 9 (i - i % 64) + 64 w + i % 64], T = ceil(V / 64);
    the padding of the last tile stays zero."""
    U = np.asarray(U, dtype=np.float64).reshape(4, 18, V)
    rows01 = U[:, :12, :]                       # (mu, component, site)
    q = quantise(rows01).transpose(0, 2, 1).reshape(4, V, 3, 4)
    w = pack_words(q).reshape(4, V, 9)          # (mu, site, word)
    T = (V + 63) // 64
    P = np.zeros((4, T * 64, 9), dtype=np.uint32)
    P[:, :V, :] = w
    return P.reshape(4, T, 64, 9).transpose(0, 1, 3, 2).reshape(-1)


def unpack_config(P, V):
    """P -> FP32 rows 0,1 in the compact layout (48 V floats), decoded as
    (float)q * (float)(1 / (2^23 - 1))."""
    T = (V + 63) // 64
    w = np.asarray(P, dtype=np.uint32).reshape(4, T, 9, 64).transpose(0, 1, 3, 2).reshape(4, T * 64, 9)[:, :V]
    q = unpack_ints(w.reshape(4, V, 3, 3)).reshape(4, V, 12)
    v = q.astype(np.float32) * np.float32(1.0 / Q)
    return v.transpose(0, 2, 1).reshape(-1)
