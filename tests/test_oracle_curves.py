"""Pins of the oracle's round Catmull-Rom curves (the primitive of the
reference's .hair strands, owlCurvesSetDegree(..., 3, forceCaps),
render_path_tracing.cu:223, render_hair_msnn.cu:346; its intersector is
closed, so the primitive is the swept circle the reference's own
surfaceNormal describes, curve_utils.cuh:159-197):

- the cubic is the uniform Catmull-Rom segment: P(0) = q1, P(1) = q2,
  P'(0) = (q2 - q0) / 2, P'(1) = (q3 - q1) / 2, which fixes a cubic;
- collinear, equally spaced control points of constant radius give a
  cylinder: the FP32 iteration and the FP64 search both meet the closed form,
  end caps included;
- the FP32 iteration (the GPU build's operations) against the FP64 search
  for the first crossing of the surface, on hair-like strands and on
  contorted ones, with bounded miss and wrong-root rates that only grazing
  rays reach;
- the hit frame: the point lies on the swept circle, and the normal is the
  gradient of the surface's distance function (finite differences in FP64).
"""
import numpy as np
import pytest

from tests import oracle_lib as ol


def seg(points, radii, strand=0, flags=0):
    q = np.zeros(20, np.float32)
    for k in range(4):
        q[4 * k:4 * k + 3] = points[k]
        q[4 * k + 3] = radii[k]
    q[16], q[17] = strand, flags
    return q


def hair_like(n, rng, r0=0.004, bend=0.15, vary=0.1):
    """Segments of hair-like strands: steps of 8-20 radii turning by ~bend rad."""
    out = np.zeros((n, 20), np.float32)
    for i in range(n):
        p = rng.normal(size=3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pts = [p]
        for _ in range(3):
            d = d + bend * rng.normal(size=3)
            d /= np.linalg.norm(d)
            pts.append(pts[-1] + d * rng.uniform(8, 20) * r0)
        out[i] = seg(pts, r0 * (1 + vary * rng.uniform(-1, 1, 4)))
    return out


def rays_at(crv, rng, spread=0.8):
    """A ray per segment through a point near a random point of its curve."""
    n = crv.shape[0]
    rays = np.zeros((n, 6))
    us = rng.uniform(0, 1, n)
    for i in range(n):
        p = ol.curve_eval(crv[i], us[i])[0]
        target = p[:3] + rng.normal(size=3) * p[3] * spread
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        rays[i, :3] = target - d * rng.uniform(0.05, 1.0)
        rays[i, 3:] = d
    return rays


def agreement(crv, rays, flags=0):
    out = ol.curve_pairs(crv, rays, flags)
    r = crv[:, [3, 7, 11, 15]].max(1)
    h32, h64 = out["hit32"], out["hit64"]
    both = h32 & h64
    pen = -out["clear"] / r  # how deep the ray enters the tube, in radii
    wrong = both & (np.abs(out["t32"].astype(np.float64) - out["t64"]) > 0.01 * r)
    return out, r, h32, h64, both, pen, wrong


@pytest.mark.parametrize("prec", [0, 1])
def test_cubic_is_the_catmull_rom_segment(prec):
    rng = np.random.default_rng(1)
    for _ in range(20):
        pts = rng.normal(size=(4, 3))
        rad = rng.uniform(0.01, 0.1, 4)
        q = seg(pts, rad)
        qq = q[:16].reshape(4, 4).astype(np.float64)
        e = ol.curve_eval(q, [0.0, 1.0], prec)
        tol = 1e-12 if prec == 0 else 3e-6
        np.testing.assert_allclose(e[0, :4], qq[1], atol=tol)
        np.testing.assert_allclose(e[1, :4], qq[2], atol=tol)
        np.testing.assert_allclose(e[0, 4:], (qq[2] - qq[0]) / 2, atol=tol)
        np.testing.assert_allclose(e[1, 4:], (qq[3] - qq[1]) / 2, atol=tol)


def cylinder_hit(o, d, a, b, r):
    """FP64 ray / cylinder around a->b: the entry into the infinite cylinder
    (its exit when the origin is inside), a hit when its axial s lies in
    [0, 1]. A ray entering past an end and leaving through the wall does not
    hit: the strand continues past the end in the next segment."""
    ax = b - a
    L = np.linalg.norm(ax)
    w = ax / L
    oo, dd = o - a, d
    op, dp = oo - (oo @ w) * w, dd - (dd @ w) * w
    A, B, C = dp @ dp, 2 * op @ dp, op @ op - r * r
    disc = B * B - 4 * A * C
    if A == 0 or disc < 0:
        return None
    t0, t1 = sorted([(-B - np.sqrt(disc)) / (2 * A), (-B + np.sqrt(disc)) / (2 * A)])
    t = t0 if t0 > 0 else t1
    if not t > 0:
        return None
    s = (oo + t * dd) @ w / L
    return (t, s) if 0 <= s <= 1 else None


def test_straight_tube_is_a_cylinder():
    rng = np.random.default_rng(2)
    a = np.array([0.1, -0.2, 0.3])
    e = np.array([0.3, 0.1, -0.2])
    r = 0.02
    q = seg([a - e, a, a + e, a + 2 * e], [r] * 4)
    hits = 0
    for i in range(400):
        s = rng.uniform(-0.1, 1.1)
        target = a + s * e + rng.normal(size=3) * r
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = target - d * rng.uniform(0.1, 1.0)
        ray = np.concatenate([o, d])[None]
        out = ol.curve_pairs(q[None], ray, 0)
        of, df = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
        ref = cylinder_hit(of, df, a.astype(np.float32).astype(np.float64),
                           (a + e).astype(np.float32).astype(np.float64), np.float32(r))
        assert bool(out["hit64"][0]) == (ref is not None), i
        if ref is None:
            assert not out["hit32"][0], i
            continue
        hits += 1
        t, u = ref
        assert out["t64"][0] == pytest.approx(t, abs=1e-7)
        assert out["u64"][0] == pytest.approx(u, abs=1e-6)
        assert out["hit32"][0], i
        assert out["t32"][0] == pytest.approx(t, abs=3e-6)
        assert out["u32"][0] == pytest.approx(u, abs=3e-5)
    assert hits > 100


def test_end_caps_close_the_strand():
    a = np.array([0.0, 0.0, 0.0])
    e = np.array([0.0, 0.0, 0.5])
    r = 0.05
    q = seg([a - e, a, a + e, a + 2 * e], [r] * 4)
    # along the axis from below: the u = 0 cap at distance 1 with a cap
    # there; without it the open end lets the ray in, and it meets only the
    # far cap (at z = 0.5) or nothing
    ray = np.array([[0.01, -0.02, -1.0, 0, 0, 1.0]])
    for flags, want in ((0, None), (1, (1.0, 0.0)), (2, (1.5, 1.0)), (3, (1.0, 0.0))):
        out = ol.curve_pairs(q[None], ray, flags)
        if want is None:
            assert not out["hit32"][0] and not out["hit64"][0], flags
        else:
            assert out["hit32"][0] and out["hit64"][0], flags
            assert out["t32"][0] == pytest.approx(want[0], abs=1e-6) and out["u32"][0] == want[1]
            assert out["t64"][0] == pytest.approx(want[0], abs=1e-9) and out["u64"][0] == want[1]
    # from above: the u = 1 cap at z = 0.5
    ray = np.array([[0.01, 0.02, 2.0, 0, 0, -1.0]])
    out = ol.curve_pairs(q[None], ray, 2)
    assert out["hit32"][0] and out["t32"][0] == pytest.approx(1.5, abs=1e-6) and out["u32"][0] == 1.0
    out = ol.curve_pairs(q[None], ray, 1)  # in through the open top, out through the bottom disk
    assert out["hit32"][0] and out["t32"][0] == pytest.approx(2.0, abs=1e-6) and out["u32"][0] == 0.0
    assert out["hit64"][0] and out["t64"][0] == pytest.approx(2.0, abs=1e-9)
    out = ol.curve_pairs(q[None], ray, 0)
    assert not out["hit32"][0] and not out["hit64"][0]


def test_iteration_meets_the_surface_on_hair_like_strands():
    rng = np.random.default_rng(3)
    crv = hair_like(6000, rng)
    rays = rays_at(crv, rng)
    out, r, h32, h64, both, pen, wrong = agreement(crv, rays)
    miss = h64 & ~h32
    extra = h32 & ~h64
    print(f"hair-like: hits {h64.sum()}, missed {miss.sum()}, extra {extra.sum()}, wrong root {wrong.sum()}, "
          f"max |dt|/r {np.max(np.abs(out['t32'][both] - out['t64'][both]) / r[both]):.2e}")
    assert h64.mean() > 0.3
    assert miss.sum() <= 0.001 * h64.sum() and extra.sum() <= 0.001 * h64.sum()
    # what disagrees is a grazing ray: it dips less than 1% of the radius in
    assert np.all(pen[miss | extra] < 0.01)
    assert wrong.sum() == 0
    dt = np.abs(out["t32"][both] - out["t64"][both]) / r[both]
    du = np.abs(out["u32"][both] - out["u64"][both])
    assert np.percentile(dt, 99.9) < 1e-3 and np.percentile(du, 99.9) < 1e-4


def test_iteration_meets_the_surface_on_contorted_strands():
    """Strongly bent segments (the direction turns by ~0.5 rad per control
    point, the radius varies by 30%): rays that run along the fibre one radius
    from its axis can converge to a later entry or none; bounded here."""
    rng = np.random.default_rng(4)
    crv = hair_like(8000, rng, r0=0.01, bend=0.5, vary=0.3)
    rays = rays_at(crv, rng)
    out, r, h32, h64, both, pen, wrong = agreement(crv, rays)
    miss = h64 & ~h32
    print(f"contorted: hits {h64.sum()}, missed {miss.sum()} (deeper than 10% r: {(miss & (pen > 0.1)).sum()}), "
          f"extra {(h32 & ~h64).sum()}, wrong root {wrong.sum()}")
    assert miss.sum() <= 0.005 * h64.sum()
    assert (miss & (pen > 0.1)).sum() <= 0.001 * h64.sum()
    assert (h32 & ~h64).sum() <= 0.001 * h64.sum()
    assert wrong.sum() <= 0.002 * h64.sum()


def distance_gradient(q, X, h=1e-6):
    """Gradient of F(X) = |X - P(u*)| - r(u*), u* the centre-line foot of X,
    by central differences in FP64: the swept surface's outward normal."""
    def F(x):
        us = np.linspace(-0.25, 1.25, 801)
        e = ol.curve_eval(q, us)
        k = np.argmin(np.sum((e[:, :3] - x) ** 2, 1))
        u = us[k]
        for _ in range(40):  # Newton on (x - P) . P' = 0
            ee = ol.curve_eval(q, [u, u + 1e-7])
            p, v = ee[0, :3], ee[0, 4:7]
            acc = (ee[1, 4:7] - v) / 1e-7
            g1, g2 = -(x - p) @ v, v @ v - (x - p) @ acc
            u -= g1 / g2
        e = ol.curve_eval(q, [u])[0]
        return np.linalg.norm(x - e[:3]) - e[3]
    g = np.array([(F(X + h * ax) - F(X - h * ax)) / (2 * h) for ax in np.eye(3)])
    return g / np.linalg.norm(g)


def test_hit_frame_on_the_swept_circle():
    rng = np.random.default_rng(5)
    crv = hair_like(300, rng, r0=0.01, bend=0.4, vary=0.4)
    rays = rays_at(crv, rng, spread=0.5)
    out = ol.curve_pairs(crv, rays, 0)
    idx = np.where(out["hit32"] & (out["u32"] > 0) & (out["u32"] < 1))[0][:40]
    assert len(idx) >= 30
    seg8 = np.zeros((crv.shape[0], 8), np.float32)  # chords: only strand ids are read
    osc = ol.OracleScene(ol.scene_info(), seg8, None, curves=crv)
    r8 = np.zeros((len(idx), 8), np.float32)
    r8[:, 0:3], r8[:, 4:7] = rays[idx, :3], rays[idx, 3:]
    g = ol.probe_hit(osc, r8, out["t32"][idx], idx.astype(np.int32), out["u32"][idx], prec=0)
    worst = 0.0
    for j, i in enumerate(idx):
        e = ol.curve_eval(crv[i], [float(out["u32"][i])])[0]
        p = g["p"][j]
        # moved onto the swept circle at u
        assert np.linalg.norm(p - e[:3]) == pytest.approx(e[3], rel=1e-6)
        assert abs((p - e[:3]) @ e[4:7]) <= 1e-9 * np.linalg.norm(e[4:7]) * e[3] + 1e-12
        assert g["radius"][j] == pytest.approx(e[3], rel=1e-9)
        np.testing.assert_allclose(g["X"][j], e[4:7] / np.linalg.norm(e[4:7]), atol=1e-9)
        nref = distance_gradient(crv[i], p)
        worst = max(worst, float(np.arccos(np.clip(g["n"][j] @ nref, -1, 1))))
    print(f"largest angle between the hit normal and the distance gradient: {worst:.2e} rad")
    assert worst < 1e-4
