"""Image output on the CPU (no GPU needed): the byte conversion of
Image::gammaCorrect + Image::writePNG (src/Utils/ImagePNG.cpp:64-72, :268-330)
on known values, the PNG file decoded independently here (chunk CRCs, IHDR,
zlib stream, filter bytes) back to the same pixels, and the headless
renderer's argument handling (src/Main.cpp:54-83), which ends before any
device work."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest


def decode_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_rgb8_known_values(hr):
    from happyray_amd import image
    v = np.array([0.0, 1.0, 0.5, 2.0, -1.0, np.nan, 0.25, 1e-9, 0.999999], np.float32)
    fb = np.zeros((3, 3), np.float32)
    fb.reshape(-1)[:] = v
    b = image.to_rgb8(fb, 3, 1, gamma=2.2).reshape(-1)
    # 0.5^(1/2.2) = 0.72974 -> 186.08; 0.25^(1/2.2) = 0.53252 -> 135.79
    # 1e-9^(1/2.2) = 8.2e-5 -> 0.02; 0.999999^(1/2.2) * 255 = 254.9999 -> 254
    np.testing.assert_array_equal(b, [0, 255, 186, 255, 0, 0, 135, 0, 254])
    lin = image.to_rgb8(fb, 3, 1, gamma=0).reshape(-1)   # no gamma: v * 255 truncated
    np.testing.assert_array_equal(lin, [0, 255, 127, 255, 0, 0, 63, 0, 254])


def test_png_round_trip(hr, tmp_path):
    from happyray_amd import image
    rng = np.random.default_rng(4)
    for w, h in [(1, 1), (7, 3), (64, 48)]:
        px = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        p = tmp_path / f"t{w}x{h}.png"
        image.write_png(p, px)
        np.testing.assert_array_equal(decode_png(p), px)


def test_screenshot_is_gamma_then_png(hr, tmp_path):
    from happyray_amd import image
    rng = np.random.default_rng(5)
    fb = rng.uniform(-0.2, 1.3, (40 * 30, 3)).astype(np.float32)
    p = tmp_path / "shot.png"
    image.write_screenshot(p, fb, 40, 30)
    np.testing.assert_array_equal(decode_png(p), image.to_rgb8(fb, 40, 30, 2.2))


def test_write_png_reports_io_errors(hr, tmp_path):
    from happyray_amd import HappyRayError, image
    with pytest.raises(HappyRayError):
        image.write_png(tmp_path / "no" / "such" / "dir.png", np.zeros((2, 2, 3), np.uint8))


def test_cli_arguments_without_device(hr):
    from happyray_amd import image
    assert os.path.exists(image.CLI_PATH)
    r = subprocess.run([image.CLI_PATH, "a", "b", "c"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Invalid arguments" in r.stderr
    r = subprocess.run([image.CLI_PATH, "scene.cfg"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "-nw" in r.stderr
    r = subprocess.run([image.CLI_PATH, "/nonexistent/scene.cfg", "-nw"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "loading the scene" in r.stderr
