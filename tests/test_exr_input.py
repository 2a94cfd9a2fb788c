"""OpenEXR input for environment maps (loadEnvTexture, model.cpp:158-180,
which reads with tinyexr's LoadEXR). The files are written here from the
OpenEXR layout (header attributes, offset table, scanline chunks) with each
supported compression: NONE, RLE and ZIP / ZIPS (zlib after the format's
byte predictor and even/odd split), HALF and FLOAT channels, RGBA, RGB and
a lone Y channel; the reader must return the written values exactly."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest

LINES = {0: 1, 1: 1, 2: 1, 3: 16}


def _attr(name, typ, val):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(val)) + val


def _predict_split(raw: bytes) -> bytes:
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]])  # even bytes, then odd
    d = t.astype(np.int32).copy()
    d[1:] = (t[1:].astype(np.int32) - t[:-1].astype(np.int32) + 128 + 256) % 256
    return d.astype(np.uint8).tobytes()


def _rle(data: bytes) -> bytes:
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        j = i
        while j < n and j - i < 127 and data[j] == data[i]:
            j += 1
        if j - i >= 3:
            out += bytes([j - i - 1, data[i]])
            i = j
        else:
            k = i
            while k < n and k - i < 127 and not (k + 2 < n and data[k] == data[k + 1] == data[k + 2]):
                k += 1
            out += struct.pack("b", -(k - i)) + data[i:k]
            i = k
    return bytes(out)


def write_exr(path, planes: dict, compression: int, half: bool, tiled_flag=False):
    """planes: channel name -> H x W array (file row order)."""
    names = sorted(planes)
    H, W = planes[names[0]].shape
    typ = 1 if half else 2
    ch = b"".join(n.encode() + b"\0" + struct.pack("<i", typ) + b"\0\0\0\0" + struct.pack("<ii", 1, 1) for n in names)
    hdr = struct.pack("<II", 20000630, 2 | (0x200 if tiled_flag else 0))
    hdr += _attr("channels", "chlist", ch + b"\0")
    hdr += _attr("compression", "compression", bytes([compression]))
    box = struct.pack("<iiii", 0, 0, W - 1, H - 1)
    hdr += _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box)
    hdr += _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    hdr += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0"
    lines = LINES.get(compression, 32)
    chunks = []
    for y0 in range(0, H, lines):
        raw = b""
        for y in range(y0, min(H, y0 + lines)):
            for n in names:
                raw += planes[n][y].astype(np.float16 if half else np.float32).tobytes()
        if compression == 0:
            data = raw
        elif compression == 1:
            data = _rle(_predict_split(raw))
        else:
            data = zlib.compress(_predict_split(raw))
        if compression and len(data) >= len(raw):
            data = raw  # stored uncompressed, as OpenEXR does
        chunks.append(struct.pack("<iI", y0, len(data)) + data)
    off = len(hdr) + 8 * len(chunks)
    table = b""
    for c in chunks:
        table += struct.pack("<Q", off)
        off += len(c)
    path.write_bytes(hdr + table + b"".join(chunks))


def load(path):
    import hairmsnn_amd as hm
    lib = hm._lib.load()
    w, h = C.c_int(), C.c_int()
    hm._lib.check(lib.hmsnn_exr_size(str(path).encode(), C.byref(w), C.byref(h)))
    out = np.empty((h.value, w.value, 4), np.float32)
    hm._lib.check(lib.hmsnn_load_exr(str(path).encode(), C.c_void_p(out.ctypes.data), w.value, h.value))
    return out


@pytest.mark.parametrize("compression", [0, 1, 2, 3])
@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("chans", ["RGBA", "RGB", "Y"])
def test_reader_returns_written_values(tmp_path, compression, half, chans):
    rng = np.random.default_rng(compression * 10 + half)
    H, W = 37, 53
    planes = {}
    for c in chans:
        v = np.exp(rng.normal(0, 2, (H, W))).astype(np.float32)
        v[:, :9] = 0.25  # runs for RLE
        planes[c] = v.astype(np.float16).astype(np.float32) if half else v
    p = tmp_path / f"x{compression}{int(half)}{chans}.exr"
    write_exr(p, planes, compression, half)
    img = load(p)
    assert img.shape == (H, W, 4)
    if chans == "Y":
        for k in range(3):
            np.testing.assert_array_equal(img[..., k], planes["Y"])
    else:
        for k, c in enumerate("RGB"):
            np.testing.assert_array_equal(img[..., k], planes[c])
    np.testing.assert_array_equal(img[..., 3], planes["A"] if "A" in chans else 1.0)


def test_half_specials(tmp_path):
    vals = np.array([0.0, -0.0, 1.0, 65504.0, 6.1035156e-05, 5.9604645e-08, -2.5, np.inf], np.float32)
    planes = {c: np.tile(vals, (2, 1)) for c in "RGB"}
    p = tmp_path / "h.exr"
    write_exr(p, planes, 0, True)
    img = load(p)
    np.testing.assert_array_equal(img[..., 0], planes["R"])
    assert np.signbit(img[0, 1, 0])


def test_roundtrip_with_writer(tmp_path):
    import hairmsnn_amd as hm
    lib = hm._lib.load()
    rgba = np.random.default_rng(1).random((6, 10, 4), dtype=np.float32)
    p = tmp_path / "w.exr"
    hm._lib.check(lib.hmsnn_save_exr(str(p).encode(), C.c_void_p(rgba.ctypes.data), 10, 6))
    np.testing.assert_array_equal(load(p), rgba[::-1])  # the writer flips to top-row-first


def test_unsupported_files_fail_cleanly(tmp_path):
    import hairmsnn_amd as hm
    planes = {c: np.ones((4, 4), np.float32) for c in "RGB"}
    write_exr(tmp_path / "t.exr", planes, 0, False, tiled_flag=True)
    write_exr(tmp_path / "piz.exr", planes, 4, False)
    (tmp_path / "junk.exr").write_bytes(b"not an exr")
    for name, code in (("t.exr", 5), ("piz.exr", 5), ("junk.exr", 2), ("missing.exr", 2)):
        with pytest.raises(hm.HmsnnError) as e:
            load(tmp_path / name)
        assert e.value.code == code, (name, str(e.value))


def _header_len(data: bytes) -> int:
    """Byte offset of the chunk offset table: after the magic, version and
    the null-terminated attribute list."""
    pos = 8
    while data[pos] != 0:
        pos = data.index(b"\0", pos) + 1          # name
        pos = data.index(b"\0", pos) + 1          # type
        size = struct.unpack_from("<I", data, pos)[0]
        pos += 4 + size
    return pos + 1


def test_corrupt_files_fail_with_eio(tmp_path):
    """A file cut inside its offset table, chunk offsets pointing past the end
    of the file (including values near 2^64, which must not wrap the bounds
    check) and a zero-size compression attribute are read errors, never a
    black image."""
    import hairmsnn_amd as hm
    planes = {c: np.full((20, 6), 0.5, np.float32) for c in "RGB"}
    good = tmp_path / "good.exr"
    write_exr(good, planes, 0, False)
    data = good.read_bytes()
    table = _header_len(data)
    cases = {
        "cut_table.exr": data[: table + 8 * 7 + 3],
        "past_end.exr": data[:table] + struct.pack("<Q", len(data) + 100) + data[table + 8:],
        "wrap.exr": data[:table] + struct.pack("<Q", 2 ** 64 - 2) + data[table + 8:],
        "zero_comp.exr": data.replace(_attr("compression", "compression", b"\0"),
                                      _attr("compression", "compression", b"")),
    }
    assert cases["zero_comp.exr"] != data
    for name, blob in cases.items():
        (tmp_path / name).write_bytes(blob)
        with pytest.raises(hm.HmsnnError) as e:
            load(tmp_path / name)
        assert e.value.code == 2, (name, str(e.value))
    np.testing.assert_array_equal(load(good)[..., 0], planes["R"])
