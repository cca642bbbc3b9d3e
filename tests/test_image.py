"""PNG and PFM writers (no GPU): files decoded here with the standard library and compared."""
import struct
import zlib

import numpy as np

import glslppm_amd as g


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 6
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def test_png_round_trip_flips_rows(tmp_path):
    rs = np.random.default_rng(3)
    w, h = 37, 1900  # more than one 64 KiB stored block
    img = rs.integers(0, 256, (h * w, 4), dtype=np.uint8)
    path = tmp_path / "x.png"
    g.write_png(str(path), img, w, h)
    top_down = read_png(str(path))
    assert np.array_equal(top_down[::-1].reshape(-1, 4), img)


def test_pfm_layout(tmp_path):
    rs = np.random.default_rng(4)
    w, h = 5, 3
    rad = rs.random((w * h, 4), dtype=np.float32)
    path = tmp_path / "x.pfm"
    g.write_pfm(str(path), rad, w, h)
    data = open(path, "rb").read()
    head = b"PF\n5 3\n-1.0\n"
    assert data.startswith(head)
    body = np.frombuffer(data[len(head):], "<f4").reshape(w * h, 3)
    assert np.array_equal(body, rad[:, :3])


def test_writers_reject_bad_arguments(tmp_path):
    lib = g.library()
    assert lib.ppm_write_png(None, None, 1, 1) == -1
    assert lib.ppm_write_pfm(str(tmp_path / "a.pfm").encode(), None, 1, 1) == -1
    buf = np.zeros(4, np.uint8)
    assert lib.ppm_write_png(str(tmp_path / "a.png").encode(), buf.ctypes.data, 0, 1) == -1
    assert lib.ppm_write_png(b"/nonexistent-dir/a.png", buf.ctypes.data, 1, 1) == -4
