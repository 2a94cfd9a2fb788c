"""Headless CLIs and EXR output (host-only)."""
import ctypes as C
import json
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "hairmsnn_amd" / "bin"


def read_exr_uncompressed(path):
    b = Path(path).read_bytes()
    assert struct.unpack_from("<I", b, 0)[0] == 20000630
    pos = 8
    attrs = {}
    while b[pos] != 0:
        e = b.index(0, pos)
        name = b[pos:e].decode()
        e2 = b.index(0, e + 1)
        typ = b[e + 1:e2].decode()
        size = struct.unpack_from("<I", b, e2 + 1)[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    assert attrs["compression"][1] == b"\x00"
    offsets = struct.unpack_from(f"<{h}Q", b, pos)
    img = np.zeros((h, w, 4), np.float32)
    for y in range(h):
        o = offsets[y]
        yy, size = struct.unpack_from("<iI", b, o)
        data = np.frombuffer(b, np.float32, 4 * w, o + 8).reshape(4, w)  # A, B, G, R
        img[yy] = data[::-1].T  # -> R, G, B, A
    return img


def test_exr_roundtrip_with_flip(tmp_path):
    import hairmsnn_amd as hm
    lib = hm._lib.load()
    w, h = 7, 5
    rgba = np.random.default_rng(0).random((h, w, 4), dtype=np.float32)
    p = tmp_path / "x.exr"
    hm._lib.check(lib.hmsnn_save_exr(str(p).encode(), C.c_void_p(rgba.ctypes.data), w, h))
    img = read_exr_uncompressed(p)
    np.testing.assert_array_equal(img, rgba[::-1])  # file row 0 = image top


def test_cli_binaries_exist_and_fail_cleanly(tmp_path):
    for name in ("render_path_tracing", "render_hair_msnn", "render_nrc"):
        assert (BIN / name).exists(), name
    for name in ("render_path_tracing", "render_hair_msnn", "render_nrc"):
        r = subprocess.run([str(BIN / name), str(tmp_path / "missing.json")], capture_output=True, text=True)
        assert r.returncode == 255, name  # main returns -1, like the reference on a failed load
        assert "Error loading scene" in r.stderr
