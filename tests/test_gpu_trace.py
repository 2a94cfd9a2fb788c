"""GPU BVH traversal against CPU intersection: closest-hit segment ids and
any-hit bits must be bit-exact (ties resolve to the lower segment id), for
every tree builder (binned SAH, the default; radix tree; PLOC)."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as ol
from tests.gen import unit_sphere

pytestmark = pytest.mark.gpu
SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def c0_scene(gpu, **over):
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["integrator"].update(over)
    return gpu.Scene(json_text=cfg)


def make_rays(seg, n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    k = n // 3
    # (a) from outside toward the hair ball
    o = unit_sphere(rng, k) * 4.0
    tgt = unit_sphere(rng, k) * rng.uniform(0, 1.8, (k, 1))
    d = tgt - o
    rays[:k, 0:3], rays[:k, 4:7] = o, d / np.linalg.norm(d, axis=1, keepdims=True)
    # (b) from points next to segments, random directions (secondary-like)
    m = n - k
    idx = rng.integers(0, seg.shape[0], m)
    u = rng.random((m, 1))
    p = seg[idx, 0:3] * (1 - u) + seg[idx, 4:7] * u
    o = p + unit_sphere(rng, m) * 0.012
    d = unit_sphere(rng, m)
    rays[k:, 0:3], rays[k:, 4:7] = o, d
    return rays


def test_closest_hit_bit_exact(gpu):
    with c0_scene(gpu, width=128, height=128) as sc:
        seg = sc.segments()
        rays = make_rays(seg, 200_000, 5)
        prim, t, u = sc.trace_rays(rays)
    prim_c, t_c, u_c = ol.intersect(seg, rays, use_bvh=True)
    hit = prim_c >= 0
    print(f"hit fraction {hit.mean():.3f}")
    assert 0.2 < hit.mean() < 0.95
    np.testing.assert_array_equal(prim, prim_c)
    np.testing.assert_array_equal(t[hit], t_c[hit])
    np.testing.assert_array_equal(u[hit], u_c[hit])


def test_closest_hit_vs_brute_force_subset(gpu):
    with c0_scene(gpu, width=128, height=128) as sc:
        seg = sc.segments()
        rays = make_rays(seg, 6000, 17)
        prim, t, u = sc.trace_rays(rays)
    prim_b, t_b, _ = ol.intersect(seg, rays, use_bvh=False)
    np.testing.assert_array_equal(prim, prim_b)
    np.testing.assert_array_equal(t[prim_b >= 0], t_b[prim_b >= 0])


def test_any_hit_bit_exact(gpu):
    with c0_scene(gpu, width=128, height=128) as sc:
        seg = sc.segments()
        rays = make_rays(seg, 200_000, 9)
        occ = sc.trace_rays(rays, any_hit=True)
    occ_c = ol.intersect(seg, rays, any_hit=True, use_bvh=True)
    np.testing.assert_array_equal(occ, occ_c)


def test_empty_and_degenerate_rays(gpu):
    with c0_scene(gpu, width=128, height=128) as sc:
        prim, t, u = sc.trace_rays(np.zeros((0, 8), np.float32))
        assert prim.shape == (0,)
        rays = np.zeros((3, 8), np.float32)
        rays[:, 0:3] = [[10, 10, 10], [0, 0, 0], [0, -6, 0]]
        rays[:, 4:7] = [[1, 0, 0], [0, 0, 1], [0, 1, 0]]  # away, from inside, through centre
        prim, t, u = sc.trace_rays(rays)
        occ = sc.trace_rays(rays, any_hit=True)
        seg = sc.segments()
    # ray 0 points away from the ball and must miss; rays 1 (from the centre,
    # inside the root sphere, which C0 does not render) and 2 (through the
    # centre) run radially, parallel to the strands around them: whatever
    # they hit, it is the brute-force answer, and any hit agrees with it
    prim_b, t_b, u_b = ol.intersect(seg, rays, use_bvh=False)
    assert prim[0] == -1 and occ[0] == 0
    np.testing.assert_array_equal(prim, prim_b)
    hit = prim_b >= 0
    np.testing.assert_array_equal(t[hit], t_b[hit])
    np.testing.assert_array_equal(u[hit], u_b[hit])
    np.testing.assert_array_equal(occ, hit.astype(occ.dtype))


@pytest.mark.parametrize("name,radius", [("c0_hairball.json", 16), ("c1_curly.json", 8)])
def test_ploc_tree_bit_exact(gpu, name, radius):
    """The PLOC-built tree (accel.builder = "ploc") gives the same closest and
    any hits as the CPU BVH; the tree shape changes only the work done."""
    cfg = json.loads((SCENES / name).read_text())
    cfg["integrator"].update({"width": 128, "height": 128})
    cfg["accel"] = {"builder": "ploc", "ploc_radius": radius}
    cfg.pop("surface", None)  # hair only: the CPU search below has no head sphere
    with gpu.Scene(json_text=cfg) as sc:
        seg = sc.segments()
        rays = make_rays(seg, 60_000, 9)
        prim, t, u = sc.trace_rays(rays)
        occ = sc.trace_rays(rays, any_hit=True)
        depth = int(sc.info["bvh_depth"])
    prim_c, t_c, u_c = ol.intersect(seg, rays, use_bvh=True)
    hit = prim_c >= 0
    print(f"{name} PLOC r={radius}: wide depth {depth}, hit fraction {hit.mean():.3f}")
    np.testing.assert_array_equal(prim, prim_c)
    np.testing.assert_array_equal(t[hit], t_c[hit])
    occ_c = ol.intersect(seg, rays, any_hit=True, use_bvh=True)
    np.testing.assert_array_equal(occ, occ_c)


@pytest.mark.parametrize("name,max_leaf,split", [("c0_hairball.json", 1, 1), ("c1_curly.json", 1, 1),
                                                 ("c0_hairball.json", 3, 1), ("c1_curly.json", 8, 1),
                                                 ("c1_curly.json", 1, 2), ("c0_hairball.json", 2, 4)])
def test_sah_tree_bit_exact(gpu, name, max_leaf, split):
    """The binned-SAH tree built on the host (accel.builder = "sah") gives the
    same closest and any hits as the CPU BVH, with one primitive per leaf,
    with leaves that are runs of up to max_leaf primitives, and with each
    segment's box split into `split` pieces along its length (every piece's
    leaf holding the whole segment)."""
    cfg = json.loads((SCENES / name).read_text())
    cfg["integrator"].update({"width": 128, "height": 128})
    cfg["accel"] = {"builder": "sah", "max_leaf": max_leaf, "split": split}
    cfg.pop("surface", None)
    if "procedural" in cfg["hair"]:
        cfg["hair"]["procedural"]["strands"] = min(cfg["hair"]["procedural"]["strands"], 5000)
    with gpu.Scene(json_text=cfg) as sc:
        seg = sc.segments()
        rays = make_rays(seg, 60_000, 11)
        prim, t, u = sc.trace_rays(rays)
        occ = sc.trace_rays(rays, any_hit=True)
        depth = int(sc.info["bvh_depth"])
        nseg = int(sc.info["num_segments"])
    prim_c, t_c, u_c = ol.intersect(seg, rays, use_bvh=True)
    hit = prim_c >= 0
    print(f"{name} SAH, runs of up to {max_leaf}, {split} boxes per segment: wide depth {depth}, "
          f"hit fraction {hit.mean():.3f}")
    assert seg.shape[0] == nseg
    assert hit.mean() > 0.1
    np.testing.assert_array_equal(prim, prim_c)
    np.testing.assert_array_equal(t[hit], t_c[hit])
    np.testing.assert_array_equal(u[hit], u_c[hit])
    occ_c = ol.intersect(seg, rays, any_hit=True, use_bvh=True)
    np.testing.assert_array_equal(occ, occ_c)


@pytest.mark.parametrize("builder", ["lbvh", "sah"])
def test_full_c1_bit_exact(gpu, builder):
    """The radix-tree and SAH builders on the full C1 geometry, 50,000
    strands x 64 segments = 3.2M segments: closest and any hits bit-exact
    against the CPU BVH."""
    cfg = json.loads((SCENES / "c1_curly.json").read_text())
    cfg["integrator"].update({"width": 128, "height": 128})
    cfg.pop("surface", None)  # hair only: the CPU search below has no head sphere
    cfg["accel"] = {"builder": builder}
    with gpu.Scene(json_text=cfg) as sc:
        seg = sc.segments()
        assert seg.shape[0] == 3_200_000
        rays = make_rays(seg, 60_000, 21)
        prim, t, u = sc.trace_rays(rays)
        occ = sc.trace_rays(rays, any_hit=True)
    prim_c, t_c, u_c = ol.intersect(seg, rays, use_bvh=True)
    hit = prim_c >= 0
    print(f"C1 {builder}: hit fraction {hit.mean():.3f}")
    assert 0.2 < hit.mean() < 0.99
    np.testing.assert_array_equal(prim, prim_c)
    np.testing.assert_array_equal(t[hit], t_c[hit])
    np.testing.assert_array_equal(u[hit], u_c[hit])
    occ_c = ol.intersect(seg, rays, any_hit=True, use_bvh=True)
    np.testing.assert_array_equal(occ, occ_c)
