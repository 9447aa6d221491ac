"""A LIME writer restated from the format c-lime writes (lime_binary_header.h,
lime_writer.c:262-318 of the reference's vendored c-lime): a 144-byte header
-- big-endian magic 0x456789ab, big-endian version 1, MB (0x80) and ME (0x40)
in byte 6, big-endian 64-bit data length, NUL-padded 128-byte type -- then the
data, zero-padded to a multiple of 8 bytes.  Test infrastructure: it builds
template files for the ILDG tests when the reference's c-lime is not built."""
import struct


def record(rtype: str, mb: bool, me: bool, data: bytes) -> bytes:
    flags = (0x80 if mb else 0) | (0x40 if me else 0)
    head = struct.pack(">IHBBQ", 0x456789AB, 1, flags, 0, len(data)) + rtype.encode()[:127].ljust(128, b"\0")
    return head + data + b"\0" * ((8 - len(data) % 8) % 8)


def message(records) -> bytes:
    """records: [(type, mb, me, bytes)]"""
    return b"".join(record(*r) for r in records)


def parse(buf: bytes):
    out, pos = [], 0
    while pos < len(buf):
        magic, version, flags, _, n = struct.unpack(">IHBBQ", buf[pos:pos + 16])
        assert magic == 0x456789AB and version == 1
        rtype = buf[pos + 16:pos + 144].split(b"\0")[0].decode()
        pos += 144
        out.append((rtype, bool(flags & 0x80), bool(flags & 0x40), buf[pos:pos + n]))
        pos += n + (8 - n % 8) % 8
    return out
