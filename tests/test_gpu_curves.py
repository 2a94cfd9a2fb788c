"""Round Catmull-Rom hair on the GPU (.hair strands with hair.primitive
"curve", the reference's primitive: owlCurvesSetDegree(..., 3, forceCaps),
render_path_tracing.cu:223, render_hair_msnn.cu:346):

- traversal: closest hits (segment, t, u) and any-hit bits bit-exact
  against the oracle's restatement of the same FP32 iteration, over its CPU
  BVH and by brute force;
- plain path tracing (end caps on) and the HairMSNN truncated pass (caps
  off) against the oracle per pixel.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as ol
from tests.gen import unit_sphere
from tests.hairfile import write_hair

pytestmark = pytest.mark.gpu
SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def strands(n, seed):
    """Curly strands grown outward from the unit sphere (8 segments each),
    so that some strand ends face the camera and their caps are seen."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        root = rng.normal(size=3)
        root /= np.linalg.norm(root)
        t = np.linspace(0, 1, 9)[:, None]
        curl = 0.05 * np.stack([np.cos(9 * t[:, 0] + k), np.sin(9 * t[:, 0] + k), np.zeros(9)], -1)
        out.append(root * (1 + 0.6 * t) + curl)
    return out


def scene(tmp_path, w=64, h=64, path_v2=1, n=400, records=None):
    p = tmp_path / "curly.hair"
    write_hair(p, strands(n, 5), d_thickness=0.03)
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["integrator"].update({"width": w, "height": h, "path_v2": path_v2, "spp": 1})
    cfg["hair"] = {"geometry": str(p), "sigma_a": [0.06, 0.1, 0.2], "beta_m": 0.3, "beta_n": 0.3, "alpha": 2.0}
    if records:
        cfg["hairmsnn"] = {"train_records": records, "warmup_steps": 4}
    return cfg


def make_rays(seg, n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    k = n // 3
    o = unit_sphere(rng, k) * 4.0
    tgt = unit_sphere(rng, k) * rng.uniform(0, 1.8, (k, 1))
    d = tgt - o
    rays[:k, 0:3], rays[:k, 4:7] = o, d / np.linalg.norm(d, axis=1, keepdims=True)
    # next to the strands (a radius is 0.006), random directions
    m = n - k
    idx = rng.integers(0, seg.shape[0], m)
    u = rng.random((m, 1))
    p = seg[idx, 0:3] * (1 - u) + seg[idx, 4:7] * u
    rays[k:, 0:3] = p + unit_sphere(rng, m) * 0.012
    rays[k:, 4:7] = unit_sphere(rng, m)
    return rays


def test_curve_scene_loads_as_curves(gpu, tmp_path):
    with gpu.Scene(json_text=scene(tmp_path)) as sc:
        assert int(sc.info["num_curves"]) == 400 * 8 == int(sc.info["num_segments"])
        crv = sc.curves()
    flags = crv[:, 17].reshape(400, 8)
    assert (flags[:, 0] == 1).all() and (flags[:, -1] == 2).all() and (flags[:, 1:-1] == 0).all()


def test_curve_traversal_bit_exact(gpu, tmp_path):
    with gpu.Scene(json_text=scene(tmp_path)) as sc:
        seg, crv = sc.segments(), sc.curves()
        rays = make_rays(seg, 120_000, 5)
        prim, t, u = sc.trace_rays(rays)
        occ = sc.trace_rays(rays, any_hit=True)
    prim_c, t_c, u_c = ol.curve_intersect(crv, rays, use_bvh=True, caps=True)
    hit = prim_c >= 0
    print(f"hit fraction {hit.mean():.3f}, cap hits {np.mean((u_c[hit] == 0) | (u_c[hit] == 1)):.4f}")
    assert 0.2 < hit.mean() < 0.95
    np.testing.assert_array_equal(prim, prim_c)
    np.testing.assert_array_equal(t[hit], t_c[hit])
    np.testing.assert_array_equal(u[hit], u_c[hit])
    occ_c, _, _ = ol.curve_intersect(crv, rays, any_hit=True, use_bvh=True, caps=True)
    np.testing.assert_array_equal(occ, occ_c)
    # the BVH finds what a brute-force scan finds
    sub = rays[:4000]
    prim_b, t_b, _ = ol.curve_intersect(crv, sub, use_bvh=False, caps=True)
    np.testing.assert_array_equal(prim[:4000], prim_b)
    np.testing.assert_array_equal(t[:4000][prim_b >= 0], t_b[prim_b >= 0])


@pytest.mark.parametrize("path_v2,need", [(1, 0.97), (40, 0.75)])
def test_curve_scene_path_tracing_matches_oracle(gpu, tmp_path, path_v2, need):
    with gpu.Scene(json_text=scene(tmp_path, path_v2=path_v2)) as sc:
        sc.pt_reset()
        sc.pt_passes(1)
        g = sc.last_pass()[..., :3].astype(np.float64)
        osc = ol.OracleScene(sc.info, sc.segments(), sc.env() if sc.info["has_env"] else None, curves=sc.curves())
    o, _, hit = osc.render(0, 0, 64, 64, spp=1, accum_id0=0, prec=1)
    agree = (np.abs(g - o) <= 1e-3 * np.maximum(np.abs(o), 1e-2)).all(-1).mean()
    print(f"curves, path_v2 {path_v2}: per-pixel agreement {agree:.4f}, hair pixels {(hit & 1).mean():.3f}")
    assert (hit & 1).mean() > 0.05
    assert agree >= need


def test_curve_scene_msnn_truncated_pass_matches_oracle(gpu, tmp_path):
    """HairMSNN traces the curves without end caps (render_hair_msnn.cu:346)."""
    with gpu.Scene(json_text=scene(tmp_path, path_v2=40, records=128)) as sc:
        sc.msnn_begin(BETA=1)
        sc.msnn_frames(1)
        short = sc.msnn_image(1)[..., :3].astype(np.float64)
        osc = ol.OracleScene(sc.info, sc.segments(), sc.env() if sc.info["has_env"] else None, curves=sc.curves())
    o_short, _, hit = osc.render(0, 0, 64, 64, spp=1, accum_id0=0, mode=1, kbeta=0, prec=1)
    o_full, _, _ = osc.render(0, 0, 64, 64, spp=1, accum_id0=0, mode=2, kbeta=0, prec=1)
    close_short = (np.abs(short - o_short) <= 1e-3 * np.maximum(np.abs(o_short), 1e-2)).all(-1)
    close_full = (np.abs(short - o_full) <= 1e-3 * np.maximum(np.abs(o_full), 1e-2)).all(-1)
    frac = (close_short | close_full).mean()
    print(f"curves, HairMSNN truncated pass: agreement {frac:.4f}, hair pixels {(hit & 1).mean():.3f}")
    assert (hit & 1).mean() > 0.05
    assert frac >= 0.95
