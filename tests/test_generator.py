"""Pins the seeded procedural hair generator (hmsnn_generate_host, host only)
to the geometry its configuration describes: roots uniform on the root
sphere; C1's curly strands are helices of the given radius and pitch with
the given segments per turn, grown along the root normal; C0's strands are
random walks of fixed segment length whose direction takes a Gaussian curl
of the given sigma per segment (checked against an independent simulation
of that walk). Each check fails for a plausible slip (radius vs diameter,
pitch per segment vs per turn, sigma vs variance)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

ROOT = Path(__file__).resolve().parent.parent
SCENES = ROOT / "hairmsnn_amd" / "scenes"


def generate(name, strands):
    import hairmsnn_amd as hm
    cfg = json.loads((SCENES / name).read_text())
    cfg["hair"]["procedural"]["strands"] = strands
    seg, _ = hm.generate_host(json.dumps(cfg))
    p = cfg["hair"]["procedural"]
    return seg.astype(np.float64), p


def strands_points(seg, nseg):
    s = seg.reshape(-1, nseg, 8)
    pts = np.concatenate([s[:, :, 0:3], s[:, -1:, 4:7]], 1)  # (strands, nseg + 1, 3)
    # consecutive segments share their end points
    np.testing.assert_array_equal(s[:, 1:, 0:3], s[:, :-1, 4:7])
    return s, pts


def test_roots_uniform_on_the_root_sphere():
    seg, p = generate("c1_curly.json", 4000)
    s, pts = strands_points(seg, p["segments"])
    root = pts[:, 0]
    np.testing.assert_allclose(np.linalg.norm(root, axis=1), p["root_sphere_radius"], rtol=1e-6)
    # z uniform in [-1, 1] and the azimuth uniform: area-uniform on the sphere
    assert stats.kstest((root[:, 2] + 1) / 2, "uniform").pvalue > 1e-3
    az = (np.arctan2(root[:, 1], root[:, 0]) + np.pi) / (2 * np.pi)
    assert stats.kstest(az, "uniform").pvalue > 1e-3
    # segment ids and radius
    np.testing.assert_array_equal(s[:, :, 7], np.repeat(np.arange(s.shape[0])[:, None], p["segments"], 1))
    np.testing.assert_allclose(s[:, :, 3], p["radius"], rtol=1e-7)


def test_curly_strands_are_helices():
    seg, p = generate("c1_curly.json", 2000)
    _, pts = strands_points(seg, p["segments"])
    n = pts[:, 0] / np.linalg.norm(pts[:, 0], axis=1, keepdims=True)
    rel = pts - pts[:, :1]
    k = np.arange(p["segments"] + 1)
    # rise along the root normal: pitch per turn of segments_per_turn segments
    rise = np.einsum("skc,sc->sk", rel, n)
    np.testing.assert_allclose(rise, np.broadcast_to(p["helix_pitch"] * k / p["segments_per_turn"], rise.shape),
                               atol=3e-6)
    # the perpendicular part moves on a circle of radius helix_radius: chords
    # between points j apart are 2 R sin(pi j / segments_per_turn), and a
    # full turn returns to the same offset
    perp = rel - rise[..., None] * n[:, None, :]
    R, T = p["helix_radius"], p["segments_per_turn"]
    for j in (1, 3, T // 2):
        chord = np.linalg.norm(perp[:, j:] - perp[:, :-j], axis=-1)
        np.testing.assert_allclose(chord, 2 * R * math.sin(math.pi * j / T), atol=3e-6)
    np.testing.assert_allclose(perp[:, T:], perp[:, :-T], atol=3e-6)
    # the root lies on the circle: helix_radius from the centre, the mean of
    # whole turns of equally spaced points
    assert p["segments"] % T == 0
    r_root = np.linalg.norm(perp[:, 0] - perp[:, :-1].mean(1), axis=-1)
    np.testing.assert_allclose(r_root, R, atol=3e-6)


def test_ball_strands_random_walk():
    seg, p = generate("c0_hairball.json", 4000)
    s, pts = strands_points(seg, p["segments"])
    d = s[:, :, 4:7] - s[:, :, 0:3]
    L = np.linalg.norm(d, axis=-1)
    np.testing.assert_allclose(L, p["segment_length"], atol=3e-7)  # FP32 end points
    u = d / L[..., None]
    n = pts[:, 0] / np.linalg.norm(pts[:, 0], axis=1, keepdims=True)
    # the first segment is the normal plus one curl step, later ones one
    # step from the previous: compare the squared step lengths with an
    # independent simulation of normalize(d + sigma g)
    sig = p["curl_sigma"]
    steps = np.concatenate([np.sum((u[:, :1] - n[:, None]) ** 2, -1), np.sum((u[:, 1:] - u[:, :-1]) ** 2, -1)], 1)
    rng = np.random.default_rng(0)
    m = 400_000
    d0 = np.tile([0.0, 0.0, 1.0], (m, 1))
    d1 = d0 + sig * rng.normal(size=(m, 3))
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    sim = np.sum((d1 - d0) ** 2, 1)
    print(f"mean squared curl step: generator {steps.mean():.6f}, simulation {sim.mean():.6f}")
    assert steps.mean() == pytest.approx(sim.mean(), rel=0.02)
    assert stats.ks_2samp(steps.ravel()[:200_000], sim[:200_000]).pvalue > 1e-3
