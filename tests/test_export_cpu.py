"""Frame export (rc_encode_png, host only): the PNG writer of rc_export_frame
checked against an independent decoder (Python's zlib and struct): signature,
IHDR fields, every chunk's CRC-32, the zlib stream, filter bytes and the pixel
rows, including the bottom-up -> top-down flip FrameExporter applies
(src/main/java/com/retrocam/export/FrameExporter.java:37-43)."""
import ctypes as C
import struct
import zlib

import numpy as np

from retrocam_amd import _native as N


def decode_png(data: bytes):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        chunks.append((kind, body))
        pos += 12 + n
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, lace = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, lace) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)  # filter type 0 on every row
    return rows[:, 1:].reshape(h, w, 4)


def test_png_round_trip_and_flip(tmp_path):
    lib = N.hip_lib()
    rng = np.random.default_rng(5)
    for w, h in ((37, 23), (300, 250)):  # the second needs several stored deflate blocks
        img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)  # rows bottom-up, as in GL
        p = str(tmp_path / f"a_{w}.png")
        assert lib.rc_encode_png(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, 1, p.encode()) == N.RC_OK
        got = decode_png(open(p, "rb").read())
        assert np.array_equal(got, img[::-1])
        assert lib.rc_encode_png(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, 0, p.encode()) == N.RC_OK
        assert np.array_equal(decode_png(open(p, "rb").read()), img)


def test_png_rejects_bad_arguments(tmp_path):
    lib = N.hip_lib()
    img = np.zeros((4, 4, 4), np.uint8)
    ptr = img.ctypes.data_as(C.POINTER(C.c_uint8))
    p = str(tmp_path / "x.png").encode()
    assert lib.rc_encode_png(ptr, 0, 4, 0, p) == N.RC_ERR_INVALID_ARG
    assert lib.rc_encode_png(None, 4, 4, 0, p) == N.RC_ERR_INVALID_ARG
    assert lib.rc_encode_png(ptr, 4, 4, 0, b"/nonexistent-dir/x.png") == N.RC_ERR_INVALID_ARG
