"""Cem Yuksel .hair input (scene.cpp:10-75): strands become uniform
Catmull-Rom curves through their points, with a phantom point mirrored past
each end and radius 0.2 x thickness; as round curves (hair.primitive
"curve", the default) the segments keep their control points, as "linear"
they are split into linear pieces."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests.hairfile import catmull_rom, write_hair

SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def helix(n, r=0.1, seed=0):
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 3, n + 1)
    base = rng.normal(size=3)
    return np.stack([r * np.cos(4 * t) + base[0], r * np.sin(4 * t) + base[1], 0.3 * t + base[2]], -1)


def scene_with(path, tess):
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["hair"] = {"geometry": str(path), "primitive": "linear", "tessellation": tess, "sigma_a": [0.06, 0.1, 0.2],
                   "beta_m": 0.3, "beta_n": 0.3, "alpha": 2.0}
    return cfg


def segments(cfg):
    import hairmsnn_amd as hm
    seg, _ = hm.generate_host(cfg)
    return seg


def test_chords_with_one_piece_per_segment(tmp_path):
    strands = [helix(n, seed=n) for n in (3, 7, 1, 12)]
    p = tmp_path / "a.hair"
    write_hair(p, strands, d_thickness=0.05)
    seg = segments(scene_with(p, 1))
    assert len(seg) == sum(len(s) - 1 for s in strands)
    k = 0
    for sid, s in enumerate(strands):
        for i in range(len(s) - 1):
            np.testing.assert_allclose(seg[k, 0:3], s[i], rtol=0, atol=1e-6)
            np.testing.assert_allclose(seg[k, 4:7], s[i + 1], rtol=0, atol=1e-6)
            assert seg[k, 3] == pytest.approx(0.2 * 0.05, rel=1e-6)
            assert seg[k, 7] == sid
            k += 1


def test_tessellated_pieces_lie_on_the_catmull_rom_curve(tmp_path):
    strands = [helix(9, seed=1), helix(4, seed=2)]
    thick = [np.linspace(0.04, 0.08, len(s)) for s in strands]
    p = tmp_path / "b.hair"
    write_hair(p, strands, thickness=thick, with_colors=True)
    tess = 4
    seg = segments(scene_with(p, tess)).astype(np.float64)
    assert len(seg) == tess * sum(len(s) - 1 for s in strands)
    k = 0
    for s, th in zip(strands, thick):
        s = np.asarray(s, np.float64)
        cp = np.concatenate([[2 * s[0] - s[1]], s, [2 * s[-1] - s[-2]]])  # phantom end points
        w = 0.2 * np.concatenate([[th[0]], th, [th[-1]]])
        for i in range(len(s) - 1):
            for j in range(tess):
                a = catmull_rom(*cp[i:i + 4], j / tess)
                e = catmull_rom(*cp[i:i + 4], (j + 1) / tess)
                r = catmull_rom(*w[i:i + 4], (j + 0.5) / tess)
                np.testing.assert_allclose(seg[k, 0:3], a, atol=2e-6)
                np.testing.assert_allclose(seg[k, 4:7], e, atol=2e-6)
                assert seg[k, 3] == pytest.approx(r, rel=1e-5)
                k += 1
        # the curve interpolates the strand's points
        starts = seg[k - tess * (len(s) - 1):k:tess, 0:3]
        np.testing.assert_allclose(starts, s[:-1], atol=1e-6)


def test_round_curves_keep_the_control_points(tmp_path):
    """hair.primitive "curve" (the default): one record per Catmull-Rom
    segment with its four control points (phantoms mirrored past the strand
    ends, render_path_tracing.cu:236-244 / scene.cpp:47-66), radii 0.2 x
    thickness, the strand id and the end flags; the segment records are the
    chords between consecutive strand points."""
    import hairmsnn_amd as hm
    strands = [helix(9, seed=1), helix(1, seed=2), helix(4, seed=3)]
    thick = [np.linspace(0.04, 0.08, len(s)) for s in strands]
    p = tmp_path / "c.hair"
    write_hair(p, strands, thickness=thick)
    cfg = scene_with(p, 4)
    del cfg["hair"]["primitive"]
    crv = hm.generate_curves_host(cfg).astype(np.float64)
    seg = segments(cfg).astype(np.float64)
    assert len(crv) == len(seg) == sum(len(s) - 1 for s in strands)
    k = 0
    for sid, (s, th) in enumerate(zip(strands, thick)):
        s = np.asarray(s, np.float32).astype(np.float64)
        cp = np.concatenate([[2 * s[0] - s[1]], s, [2 * s[-1] - s[-2]]])
        w = 0.2 * np.concatenate([[th[0]], th, [th[-1]]]).astype(np.float32)
        n = len(s) - 1
        for i in range(n):
            q = crv[k, :16].reshape(4, 4)
            np.testing.assert_allclose(q[:, :3], cp[i:i + 4], atol=1e-6)
            np.testing.assert_allclose(q[:, 3], w[i:i + 4], rtol=1e-6)
            assert crv[k, 16] == sid
            assert crv[k, 17] == (1 if i == 0 else 0) + (2 if i == n - 1 else 0)
            np.testing.assert_allclose(seg[k, 0:3], s[i], atol=1e-6)
            np.testing.assert_allclose(seg[k, 4:7], s[i + 1], atol=1e-6)
            assert seg[k, 7] == sid
            k += 1
    # "linear" keeps the pieces and no curves
    cfg["hair"]["primitive"] = "linear"
    assert hm.generate_curves_host(cfg).shape[0] == 0
    cfg["hair"]["primitive"] = "cubic"
    with pytest.raises(hm.HmsnnError):
        hm.generate_curves_host(cfg)


def test_default_segment_count(tmp_path):
    strands = [helix(5, seed=k) for k in range(6)]
    p = tmp_path / "c.hair"
    write_hair(p, strands, d_segments=5, d_thickness=0.1)
    seg = segments(scene_with(p, 2))
    assert len(seg) == 6 * 5 * 2
    assert np.unique(seg[:, 7]).tolist() == list(range(6))


def test_bad_files_fail_cleanly(tmp_path):
    import hairmsnn_amd as hm
    p = tmp_path / "d.hair"
    write_hair(p, [helix(3)])
    bad_sig = tmp_path / "e.hair"
    bad_sig.write_bytes(b"HAIX" + p.read_bytes()[4:])
    short = tmp_path / "f.hair"
    short.write_bytes(p.read_bytes()[:150])
    for f, msg in ((bad_sig, "wrong signature"), (short, "Cannot read hair points"),
                   (tmp_path / "missing.hair", "Cannot open hair file")):
        with pytest.raises(hm.HmsnnError) as e:
            segments(scene_with(f, 1))
        assert e.value.code == 2 and msg in str(e.value)


def test_zero_segment_strand_keeps_its_point(tmp_path):
    """A strand with 0 segments owns one point of the file (the reference's
    point loop consumes segments + 1 points per strand, scene.cpp:24-66): the
    strands after it must read their own points, and strand ids stay the
    file's strand indices."""
    strands = [helix(3, seed=1), helix(0, seed=2), helix(4, seed=3), helix(0, seed=4), helix(2, seed=5)]
    assert [len(s) - 1 for s in strands] == [3, 0, 4, 0, 2]
    p = tmp_path / "z.hair"
    write_hair(p, strands, d_thickness=0.05)
    seg = segments(scene_with(p, 1))
    assert len(seg) == 3 + 4 + 2
    k = 0
    for sid, s in enumerate(strands):
        for i in range(len(s) - 1):
            np.testing.assert_allclose(seg[k, 0:3], s[i], rtol=0, atol=1e-6)
            np.testing.assert_allclose(seg[k, 4:7], s[i + 1], rtol=0, atol=1e-6)
            assert seg[k, 7] == sid
            k += 1
