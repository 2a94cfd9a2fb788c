"""Writer of Cem Yuksel .hair files (the format the reference loads with
cyHairFile, scene.cpp:10-112) for the tests: a 128-byte header (signature
"HAIR", hair count, point count, array flags, default segments, default
thickness, default transparency, default colour, 88 bytes of text) followed
by the optional segment (uint16 per strand), point (3 float32), thickness,
transparency and colour arrays."""
import struct

import numpy as np

SEGMENTS, POINTS, THICKNESS, TRANSPARENCY, COLORS = 1, 2, 4, 8, 16


def write_hair(path, strands, thickness=None, d_segments=None, d_thickness=0.05, with_colors=False):
    """strands: list of (n_i + 1) x 3 point arrays. thickness: optional list of
    per-point arrays. d_segments: when given, no segment array is written and
    every strand must have d_segments segments."""
    pts = np.concatenate([np.asarray(s, np.float32) for s in strands])
    arrays = POINTS
    body = b""
    if d_segments is None:
        arrays |= SEGMENTS
        body += np.array([len(s) - 1 for s in strands], np.uint16).tobytes()
    body += pts.astype(np.float32).tobytes()
    if thickness is not None:
        arrays |= THICKNESS
        body += np.concatenate([np.asarray(t, np.float32) for t in thickness]).tobytes()
    if with_colors:
        arrays |= TRANSPARENCY | COLORS
        body += np.zeros(len(pts), np.float32).tobytes() + np.ones((len(pts), 3), np.float32).tobytes()
    hdr = b"HAIR" + struct.pack("<IIIIff3f", len(strands), len(pts), arrays, d_segments or 0, d_thickness, 0.0,
                                1.0, 1.0, 1.0)
    hdr += b"written for tests".ljust(88, b"\0")
    assert len(hdr) == 128
    path.write_bytes(hdr + body)


def catmull_rom(p0, p1, p2, p3, t):
    t2, t3 = t * t, t * t * t
    return 0.5 * (2 * p1 + (p2 - p0) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (3 * p1 - p0 - 3 * p2 + p3) * t3)
