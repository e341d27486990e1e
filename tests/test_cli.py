"""The `path` command line (flags of the reference's src/main.cpp)."""
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "path_amd" / "_lib" / "path"


def read_png(path):
    data = Path(path).read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(kind + body) == crc
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 3)


def test_help_and_usage_errors():
    r = subprocess.run([str(CLI), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--output-per-step" in r.stdout
    r = subprocess.run([str(CLI)], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([str(CLI), "scene.xml", "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown option" in r.stderr


@pytest.mark.gpu
def test_render_and_per_step_images(renderer, tmp_path):
    from path_amd import scenes
    from path_amd.renderer import Scene, Settings
    xml = scenes.cornell_box(tmp_path / "scene")
    r = subprocess.run([str(CLI), str(xml), "--width", "64", "--height=48", "--samples", "8",
                        "--output", str(tmp_path / "o.png"), "--output-per-step",
                        "--n-output-steps", "4", "--no-print-config", "--disable-progress"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    img = read_png(tmp_path / "o.png")
    scene, ftw = Scene.load(xml, 64 / 48)
    bgra, _ = renderer.render(scene, ftw, 64, 48, 8, Settings(), output="BGRA")
    want = np.stack([(bgra >> 16) & 255, (bgra >> 8) & 255, bgra & 255], -1).astype(np.uint8)
    assert np.array_equal(img, want)
    for i in range(4):
        assert read_png(tmp_path / f"out_{i}.png").shape == (48, 64, 3)
    assert read_png(tmp_path / "summed_image.png").shape == (48, 64, 3)
