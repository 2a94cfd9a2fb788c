"""Pins the oracle's neural-radiance-caching G_BUFFER restatement
(nrcTracePaths, cuda/nrc.cu:135-343) on the CPU.

* c = 0 makes every path query the cache at its second vertex, so the
  radiance before the query is the first vertex's direct light: exactly the
  oracle's pathTrace with path_v2 = 1 (same random stream, bit for bit).
* With an unbounded spread threshold and every pixel an unbiased training
  pixel, no path queries the cache and the per-vertex records must
  reassemble the path's radiance: the training target of vertex 0 (the
  reference's suffix sum, which stops before the last vertex) plus the last
  vertex's direct light times the throughput to it.
* Without Russian roulette and without cache queries the G-buffer radiance
  estimates the same expectation as pathTrace (which has roulette): the two
  image means agree within their Monte Carlo errors.
"""
import json
from pathlib import Path

import numpy as np

from tests import oracle_lib as ol

SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def make_scene(w, h, path_v2=40, strands=600, furnace=False):
    import hairmsnn_amd as hm
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["integrator"].update({"width": w, "height": h, "path_v1": 1, "path_v2": path_v2})
    cfg["hair"]["procedural"]["strands"] = strands
    if furnace:
        cfg["hair"]["sigma_a"] = [0.0, 0.0, 0.0]
        cfg["lights"] = {"environment": {"constant": [1.0, 1.0, 1.0], "resolution": [64, 32], "scale": 1.0}}
    info = hm.parse_scene(json.dumps(cfg))
    seg, envmap = hm.generate_host(cfg)
    e = {"rgba": envmap, **ol.env_tables(envmap, 0)} if envmap is not None and info["has_env"] else None
    return ol.OracleScene(info, seg, e)


def test_c_zero_radiance_is_first_vertex_direct_light():
    w = h = 32
    pt = make_scene(w, h, path_v2=1)
    full, _, hit = pt.render(0, 0, w, h, spp=1, accum_id0=0, mode=0, prec=1)
    nrc = make_scene(w, h, path_v2=40)
    out = nrc.nrc_gbuffer(0, np.zeros(w * h, np.uint8), c=0.0, scene_scale=1.0, prec=1)
    rad = out["rad"].reshape(h, w, 3)
    np.testing.assert_array_equal(rad, full)
    assert (out["gb_hit"].reshape(h, w) == (hit & 1)).all()
    # every path whose continuation hit something queried the cache there, with
    # the first vertex's f cos / pdf as throughput
    q = out["query"] == 1
    assert q.sum() > 0.2 * out["gb_hit"].sum()
    assert (np.abs(out["beta"][q]).sum(-1) > 0).all()
    assert (out["beta"][~q] == 0).all()


def reference_targets(rec, bounces, cache_rgb):
    """nrcGenerateTrainingData (cuda/nrc.cu:69-133) for one training pixel:
    target of vertex b = sum over b <= sub < bounces of (product of the
    continuation factors b..sub-1) x direct light(sub), plus the product over
    b..bounces-1 times the suffix's cache value."""
    out = []
    for b in range(bounces):
        color = np.zeros(3)
        beta = np.ones(3)
        for sub in range(b, bounces):
            color = color + beta * rec[sub, 9:12]
            beta = beta * rec[sub, 12:15]
        out.append(color + beta * cache_rgb)
    return np.array(out).reshape(-1, 3)


def test_records_reassemble_the_path_radiance():
    w = h = 32
    osc = make_scene(w, h, strands=2000)
    flags = np.full(w * h, 3, np.uint8)  # every pixel a training pixel with an unbiased suffix
    out = osc.nrc_gbuffer(0, flags, c=1e30, scene_scale=2.0, prec=0)
    hits = np.flatnonzero(out["gb_hit"])
    assert len(hits) > 100
    assert (out["query"] == 0).all() and (out["tb_hit"] == 0).all()
    multi = 0
    for i in hits:
        o = out[i]
        B = int(o["bounces"])
        rec = o["rec"]
        t = reference_targets(rec, B, np.zeros(3))
        t0 = t[0] if B > 0 else np.zeros(3)
        through = np.prod(rec[:B, 12:15], axis=0) if B > 0 else np.ones(3)
        total = t0 + through * rec[B, 9:12]
        np.testing.assert_allclose(total, o["rad"], rtol=1e-12, atol=1e-14)
        multi += B >= 2
        # records hold the vertex / scene scale and unit directions
        if B > 0:
            assert np.allclose(np.linalg.norm(rec[:B, 3:6], axis=-1), 1.0, atol=1e-6)
            assert np.allclose(np.linalg.norm(rec[:B, 6:9], axis=-1), 1.0, atol=1e-5)
    print(f"{len(hits)} hit pixels, {multi} with two or more recorded vertices")
    assert multi > 20


def test_unbounded_spread_matches_path_tracer_expectation():
    w = h = 32
    frames = 16
    osc = make_scene(w, h, furnace=True)
    zeros = np.zeros(w * h, np.uint8)
    nrc = np.stack([osc.nrc_gbuffer(f, zeros, c=1e30, scene_scale=1.0, prec=0)["rad"] for f in range(frames)])
    pt = np.stack([osc.render(0, 0, w, h, spp=1, accum_id0=f, mode=0, prec=0)[0].reshape(-1, 3)
                   for f in range(frames)])
    hit = osc.nrc_gbuffer(0, zeros, c=1e30, scene_scale=1.0, prec=0)["gb_hit"] == 1
    a = nrc[:, hit].mean(-1)  # frames x hit pixels
    b = pt[:, hit].mean(-1)
    ma, mb = a.mean(), b.mean()
    se = np.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
    print(f"furnace hit pixels: NRC (no cache) {ma:.5f}  pathTrace {mb:.5f}  se {se:.5f}")
    assert abs(ma - mb) < 4 * se + 1e-3
    # the white furnace: radiance close to 1 (depth-40 truncation aside)
    assert 0.9 < ma < 1.01
