"""Pins the oracle's environment sampling, GGX surface, hit geometry and
hair frame, the hair lobes' longitudinal tilts and the reference's MIS
weighting with a directional light (quirk Q1) against what the mathematics
fixes: closed forms, normalisation integrals, chi-square tests of samplers
against their pdfs, geometric invariants and a quadrature of the direct
lighting estimator's expectation. Each check is chosen so that a plausible
slip in the oracle (a dropped term, a swapped sign or index) fails it."""
import math

import numpy as np
import pytest
from scipy import stats

from tests import oracle_lib as ol
from tests.gen import ALPHA_2DEG, unit_sphere

# FP32 constants of the reference: Pi (utils.cuh:10) and the host's 3.14159f
PI_F = float(np.float32(3.1415926))


def procedural_env(W=32, H=16, seed=7):
    """Smooth-plus-noise RGB environment with a bright spot (RGBA32F)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:H, 0:W]
    base = 0.3 + 0.2 * np.sin(2 * np.pi * x / W)[..., None] * np.array([1.0, 0.5, 0.2])
    noise = rng.lognormal(0.0, 0.5, (H, W, 3)) * 0.2
    spot = 5.0 * np.exp(-(((x - 0.7 * W) / 2.0) ** 2 + ((y - 0.35 * H) / 1.5) ** 2))[..., None]
    rgb = base + noise + spot
    return np.concatenate([rgb, np.ones((H, W, 1))], -1).astype(np.float32)


def env_scene(rgba, env_pdf=1, rot=0.0, scale=1.0, **kw):
    info = ol.scene_info(has_env=1, env_pdf=env_pdf, env_rot_phi=rot, env_scale=scale, **kw)
    seg = np.zeros((1, 8), np.float32)
    seg[0] = [-1, 0, 0, 0.05, 1, 0, 0, 0]
    return ol.OracleScene(info, seg, ol.env_with_tables(rgba))


def dirs_from_uv(u, v, rot=0.0):
    """World direction of texture coordinates (u, v): phi = 2 pi u - rot,
    theta = pi v (optix_common.cuh:137-145)."""
    th = np.pi * v
    ph = 2 * np.pi * u - rot
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)


# ---------------------------------------------------------------------------
# environment (optix_common.cuh:10-149, scene.cpp:349-425)

def test_env_radiance_texel_centres_and_bilinear_midpoints(orc):
    rgba = procedural_env()
    H, W = rgba.shape[:2]
    osc = env_scene(rgba, scale=2.0)
    ys, xs = np.mgrid[1:H - 1, 0:W - 1]
    ys, xs = ys.ravel(), xs.ravel()
    # texel centres: the texel's value times the scale
    L, _ = ol.probe_env(osc, dirs_from_uv((xs + 0.5) / W, (ys + 0.5) / H))
    np.testing.assert_allclose(L, 2.0 * rgba[ys, xs, :3], rtol=2e-5)
    # midway between two texel centres along u: their average
    L, _ = ol.probe_env(osc, dirs_from_uv((xs + 1.0) / W, (ys + 0.5) / H))
    np.testing.assert_allclose(L, rgba[ys, xs, :3] + rgba[ys, xs + 1, :3], rtol=2e-5)
    # ... and along v (rows run with theta: row 0 at the +z pole)
    L, _ = ol.probe_env(osc, dirs_from_uv((xs + 0.5) / W, (ys + 1.0) / H))
    np.testing.assert_allclose(L, rgba[ys, xs, :3] + rgba[ys + 1, xs, :3], rtol=2e-5)
    # a rotation by one texel moves the lookup one column
    osc_r = env_scene(rgba, rot=2 * np.pi / W)
    L0, _ = ol.probe_env(env_scene(rgba), dirs_from_uv((xs + 1.5) / W, (ys + 0.5) / H))
    Lr, _ = ol.probe_env(osc_r, dirs_from_uv((xs + 1.5) / W, (ys + 0.5) / H, rot=2 * np.pi / W))
    np.testing.assert_allclose(Lr, L0, rtol=2e-4)


def test_constant_env_pdf_closed_form(orc):
    """Constant radiance: p(theta) = sin(theta_row) H sin(pi / 2H) / (2 pi^2 sin theta),
    the row's centre sine over the direction's, times the normaliser of the
    rows' sines (sum_v sin(pi (v + 1/2) / H) = 1 / sin(pi / 2H)); 1 / 4 pi
    without importance sampling."""
    W, H = 48, 24
    rgba = np.full((H, W, 4), 0.5, np.float32)
    osc = env_scene(rgba)
    d = unit_sphere(np.random.default_rng(1), 20000)
    _, pdf = ol.probe_env(osc, d)
    th = np.arccos(np.clip(d[:, 2].astype(np.float32).astype(np.float64), -1, 1))
    row = np.clip((th / PI_F * H).astype(int), 0, H - 1)
    s_row = np.sin(3.14159 * (row + 0.5) / H)
    expect = s_row * H * np.sin(np.pi / (2 * H)) / (2 * PI_F ** 2 * np.sin(th))
    np.testing.assert_allclose(pdf, expect, rtol=1e-4)
    _, pdf_u = ol.probe_env(env_scene(rgba, env_pdf=0), d)
    np.testing.assert_allclose(pdf_u, 1 / (4 * PI_F), rtol=1e-7)


def test_env_pdf_integrates_to_one(orc):
    """Midpoint quadrature of p(omega) sin(theta) d theta d phi on a grid four
    times finer than the texels (p sin(theta) is constant per texel)."""
    rgba = procedural_env()
    H, W = rgba.shape[:2]
    osc = env_scene(rgba)
    k = 4
    v = (np.arange(H * k) + 0.5) / (H * k)
    u = (np.arange(W * k) + 0.5) / (W * k)
    V, U = np.meshgrid(v, u, indexing="ij")
    _, pdf = ol.probe_env(osc, dirs_from_uv(U.ravel(), V.ravel()))
    integral = np.sum(pdf * np.sin(np.pi * V.ravel())) * (np.pi / (H * k)) * (2 * np.pi / (W * k))
    assert integral == pytest.approx(1.0, abs=2e-5)


def test_env_sampling_chi_square_and_pdf_consistency(orc):
    """Texel histogram of 400k samples against the probability the
    environment itself gives each texel, mean RGB x sin(theta_row), computed
    here from the texels (not from the tables); the returned pdf is the pdf
    at the sampled direction and the returned radiance the lookup there."""
    rgba = procedural_env()
    H, W = rgba.shape[:2]
    osc = env_scene(rgba)
    n = 400000
    seeds = np.random.default_rng(3).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    wi, L, pdf = ol.probe_env_sample(osc, seeds)
    th = np.arccos(np.clip(wi[:, 2], -1, 1))
    ph = np.mod(np.arctan2(wi[:, 1], wi[:, 0]), 2 * np.pi)
    iu = np.clip((ph / (2 * np.pi) * W).astype(int), 0, W - 1)
    iv = np.clip((th / np.pi * H).astype(int), 0, H - 1)
    obs = np.bincount(iv * W + iu, minlength=W * H).astype(np.float64)
    w = rgba[..., :3].astype(np.float64).mean(-1) * np.sin(np.pi * (np.arange(H)[:, None] + 0.5) / H)
    expect = (w / w.sum()).ravel() * n
    chi2 = np.sum((obs - expect) ** 2 / expect)
    dof = W * H - 1
    assert chi2 < stats.chi2.ppf(0.9999, dof), (chi2, dof)
    # within a texel the pdf is constant in (u, v): the fractional texture
    # coordinates are uniform (Kolmogorov-Smirnov, rows away from the poles)
    fu, fv = ph / (2 * np.pi) * W, th / np.pi * H
    mid = (iv >= 2) & (iv < H - 2)
    for frac in (np.mod(fu[mid], 1.0), np.mod(fv[mid], 1.0)):
        assert stats.kstest(frac, "uniform").pvalue > 1e-4
    # sampled pdf == pdf at the sample, away from texel edges (where the FP32
    # round trip through atan2 / acos can land in the neighbour texel) and
    # from the poles (where the FP32 direction's z fixes theta poorly)
    inner = (np.abs(fu - np.round(fu)) > 1e-3) & (np.abs(fv - np.round(fv)) > 1e-3) & (np.abs(wi[:, 2]) < 0.99)
    L2, pdf2 = ol.probe_env(osc, wi)
    np.testing.assert_allclose(pdf[inner], pdf2[inner], rtol=1e-4)
    np.testing.assert_allclose(L[inner], L2[inner], rtol=1e-3)
    assert np.isfinite(pdf).all() and (pdf > 0).all()


# ---------------------------------------------------------------------------
# GGX head surface (frostbite_anisotropic.cuh:26-157)

def hemisphere_grid(n_t=512, n_p=1024):
    t = (np.arange(n_t) + 0.5) / n_t * (np.pi / 2)
    p = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, P = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n_t) * (2 * np.pi / n_p)).ravel()
    return d, dw


@pytest.mark.parametrize("a", [0.2, 0.5, 1.0])
def test_ggx_ndf_normalised(orc, a):
    """int D(h) (n . h) d omega_h = 1 over the hemisphere."""
    h, dw = hemisphere_grid()
    assert np.sum(ol.ggx_d(a, h) * h[:, 2] * dw) == pytest.approx(1.0, rel=2e-4)


def smith_lambda(a, w):
    t2 = (w[:, 0] ** 2 + w[:, 1] ** 2) / w[:, 2] ** 2
    return 0.5 * (-1 + np.sqrt(1 + a * a * t2))


@pytest.mark.parametrize("a", [0.1, 0.5, 1.0])
def test_ggx_sample_weight_closed_form(orc, a):
    """Visible-normal sampling makes the specular weight f_spec cos / pdf
    equal to 0.5 G2 / G1 = 0.5 (1 + L(o)) / (1 + L(o) + L(i)) (Heitz 2018),
    with the Smith Lambda written here from its definition; with albedo the
    diffuse half adds 0.5 albedo / pi |cos| exactly."""
    rng = np.random.default_rng(5)
    n = 20000
    ct = rng.uniform(0.05, 1.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    wo = np.stack([st * np.cos(ph), st * np.sin(ph), ct], 1)
    u2 = rng.random((n, 2))
    wi, pdf, f = ol.ggx_sample(a, wo, u2)
    up = wi[:, 2] > 1e-4
    assert up.mean() > 0.5
    weight = f[up, 0] / pdf[up]
    lo, li = smith_lambda(a, wo[up]), smith_lambda(a, wi[up])
    np.testing.assert_allclose(weight, 0.5 * (1 + lo) / (1 + lo + li), rtol=1e-9)
    # lower-hemisphere samples carry no value
    assert np.all(f[~up & (wi[:, 2] < 0)] == 0)
    fa, _ = ol.ggx_eval(a, wo[up], wi[up], albedo=(0.4, 0.4, 0.4))
    np.testing.assert_allclose(fa[:, 0] - f[up, 0], 0.5 * 0.4 / PI_F * wi[up, 2], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("a,ct", [(0.3, 0.8), (0.7, 0.3)])
def test_ggx_sampling_chi_square_and_pdf_normalised(orc, a, ct):
    """Samples follow pdf_GGX; pdf_GGX integrates to one over the directions
    whose half vector faces the viewer."""
    st = math.sqrt(1 - ct * ct)
    wo1 = np.array([st, 0.0, ct])
    n = 300000
    rng = np.random.default_rng(9)
    wi, pdf, _ = ol.ggx_sample(a, np.repeat(wo1[None], n, 0), rng.random((n, 2)))
    # full sphere grid in (cos theta, phi): d omega = d(cos) d(phi)
    nb_c, nb_p, m = 16, 32, 8
    c = ((np.arange(nb_c * m) + 0.5) / (nb_c * m)) * 2 - 1
    p = ((np.arange(nb_p * m) + 0.5) / (nb_p * m)) * 2 * np.pi - np.pi
    C, P = np.meshgrid(c, p, indexing="ij")
    s = np.sqrt(1 - C * C)
    dirs = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
    _, pe = ol.ggx_eval(a, np.repeat(wo1[None], dirs.shape[0], 0), dirs)
    hv = dirs + wo1
    hv /= np.linalg.norm(hv, axis=1, keepdims=True)
    ok = (hv[:, 2] > 0) & (hv @ wo1 > 0)
    pe = np.where(ok, pe, 0.0)
    cell = (2 / (nb_c * m)) * (2 * np.pi / (nb_p * m))
    assert np.sum(pe) * cell == pytest.approx(1.0, abs=3e-3)
    exp = (pe.reshape(nb_c, m, nb_p, m).sum((1, 3)) * cell).ravel() * n
    ci = np.clip(((wi[:, 2] + 1) / 2 * nb_c).astype(int), 0, nb_c - 1)
    pi_ = np.clip(((np.arctan2(wi[:, 1], wi[:, 0]) + np.pi) / (2 * np.pi) * nb_p).astype(int), 0, nb_p - 1)
    obs = np.bincount(ci * nb_p + pi_, minlength=nb_c * nb_p).astype(np.float64)
    keep = exp > 20
    chi2 = np.sum((obs[keep] - exp[keep]) ** 2 / exp[keep])
    dof = keep.sum() - 1
    assert chi2 < stats.chi2.ppf(0.9999, dof) + 0.003 * n * 0, (chi2, dof)
    assert obs[~keep].sum() < 0.002 * n
    # returned pdf equals pdf_GGX at the sampled direction
    _, pdf2 = ol.ggx_eval(a, np.repeat(wo1[None], n, 0), wi)
    np.testing.assert_allclose(pdf, pdf2, rtol=1e-6)


# ---------------------------------------------------------------------------
# hit geometry and the hair frame (curve_utils.cuh:159-214, optix_common.cuh:558-599)

def random_segment_hits(nseg=300, nray=4000, seed=11):
    rng = np.random.default_rng(seed)
    p0 = rng.uniform(-1, 1, (nseg, 3))
    d = unit_sphere(rng, nseg) * rng.uniform(0.05, 0.3, (nseg, 1))
    r = rng.uniform(0.004, 0.03, nseg)
    seg = np.concatenate([p0, r[:, None], p0 + d, np.arange(nseg)[:, None] % 50], 1).astype(np.float32)
    # rays aimed at random points near random segments
    k = rng.integers(0, nseg, nray)
    tgt = p0[k] + rng.uniform(0, 1, (nray, 1)) * d[k] + unit_sphere(rng, nray) * (r[k, None] * 0.9)
    o = tgt + unit_sphere(rng, nray) * 3.0
    dd = tgt - o
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    rays = np.zeros((nray, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, dd
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7].astype(np.float64), axis=1, keepdims=True)
    prim, t, u = ol.intersect(seg, rays)
    hit = prim >= 0
    return seg, rays[hit], prim[hit], t[hit], u[hit]


@pytest.mark.parametrize("prec", [0, 1])
def test_hair_hit_geometry_and_frame(orc, prec):
    seg, rays, prim, t, u = random_segment_hits()
    assert prim.size > 2000
    osc = ol.OracleScene(ol.scene_info(), seg, None)
    g = ol.probe_hit(osc, rays, t, prim, u, prec=prec)
    tol = 1e-12 if prec == 0 else 2e-5
    s = seg[prim].astype(np.float64)
    p0, p1, r = s[:, 0:3], s[:, 4:7], s[:, 3]
    axis = p1 - p0
    ax = axis / np.linalg.norm(axis, axis=1, keepdims=True)
    rel = g["p"] - p0
    # |p - axis| = r, n radial: n = (p - foot) / r, n perpendicular to the tangent
    foot = p0 + np.sum(rel * ax, 1, keepdims=True) * ax
    dist = np.linalg.norm(g["p"] - foot, axis=1)
    np.testing.assert_allclose(dist, r, rtol=20 * tol)
    np.testing.assert_allclose(g["n"], (g["p"] - foot) / r[:, None], atol=50 * tol)
    assert np.abs(np.sum(g["n"] * ax, 1)).max() < 20 * tol
    # the refined point lies on the ray up to the refinement step (radial only)
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    on_ray = o + t[:, None].astype(np.float64) * d
    assert np.linalg.norm(on_ray - g["p"], axis=1).max() < 1e-4
    # the frame: X the tangent, orthonormal and right-handed
    F = np.stack([g["X"], g["Y"], g["Z"]], 1)
    np.testing.assert_allclose(F @ F.transpose(0, 2, 1), np.broadcast_to(np.eye(3), F.shape), atol=20 * tol)
    np.testing.assert_allclose(np.linalg.det(F), 1.0, atol=20 * tol)
    np.testing.assert_allclose(g["X"], ax, atol=20 * tol)
    # wo in the frame: y = 0 exactly by construction, z > 0
    wo_l = g["wo_local"]
    assert np.abs(wo_l[:, 1]).max() < 20 * tol
    assert (wo_l[:, 2] > 0).all()
    np.testing.assert_allclose(wo_l[:, 0], np.sum(-d * ax, 1), atol=50 * tol)
    # h = dot(Y, n): Y is perpendicular to the ray and to the axis, so every
    # point of the ray has the same Y offset from the axis, r h
    # (up to the FP32 hit distance: an error of t 2^-24 along the ray turns
    # the radial direction by up to that over r)
    h_ray = np.sum((o - p0) * g["Y"], 1) / r
    err = np.abs(g["h"] - h_ray) / (t * 2.0 ** -24 / r + (0 if prec == 0 else 1e-5))
    assert np.quantile(err, 0.999) < 4 and err.max() < 10, (np.quantile(err, 0.999), err.max())
    assert (np.abs(g["h"]) <= 1 + 1e-6).all()
    np.testing.assert_allclose(g["radius"], r, rtol=1e-7)


def test_sphere_hit_geometry(orc):
    rng = np.random.default_rng(4)
    n = 2000
    c, R = np.array([0.25, -0.125, 0.375]), 0.75  # exact in FP32
    o = unit_sphere(rng, n) * 3 + c
    tgt = c + unit_sphere(rng, n) * 0.5 * R
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # nearest root of |o + t d - c| = R in FP64
    oc = o - c
    b = np.sum(oc * d, 1)
    t = -b - np.sqrt(b * b - (np.sum(oc * oc, 1) - R * R))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, d
    info = ol.scene_info(has_sphere=1, sphere=[*c, R], sphere_albedo=[0.2, 0.3, 0.4], sphere_alpha=0.25)
    osc = ol.OracleScene(info, np.zeros((1, 8), np.float32), None)
    g = ol.probe_hit(osc, rays, t.astype(np.float32), np.full(n, -2, np.int32), np.zeros(n, np.float32))
    np.testing.assert_allclose(np.linalg.norm(g["p"] - c, axis=1), R, rtol=1e-12)
    np.testing.assert_allclose(g["n"], (g["p"] - c) / R, atol=1e-12)
    F = np.stack([g["X"], g["Y"], g["Z"]], 1)
    np.testing.assert_allclose(F @ F.transpose(0, 2, 1), np.broadcast_to(np.eye(3), F.shape), atol=1e-12)
    np.testing.assert_allclose(g["Z"], g["n"], atol=1e-12)
    np.testing.assert_allclose(g["wo_local"][:, 2], np.sum(-d * g["n"], 1), atol=1e-7)
    assert (g["wo_local"][:, 2] > 0).all()


# ---------------------------------------------------------------------------
# hair lobe tilts (disney_hair.cuh:228-254)

def lobe_curve(so, alpha, lobe, th):
    """One lobe (the gains isolate it) as a function of theta_i at phi_i = 0,
    phi_o = pi / 2, h = 0."""
    gains = [0, 0, 0, 0]
    gains[lobe] = 1
    params = ol.hair_params(0.3, 0.3, alpha, gains)
    n = th.size
    q = np.zeros((n, 10), np.float32)
    q[:, 0:3] = [so, 0.0, math.sqrt(1 - so * so)]
    q[:, 3:6] = np.stack([np.sin(th), np.cos(th), np.zeros(n)], 1)
    f, _ = ol.hair_eval(q, params, prec=0)
    return f[:, 0]


@pytest.mark.parametrize("so", [-0.4, 0.0, 0.5])
def test_lobe_tilts(orc, so):
    """The M_p of the R, TT and TRT lobes use theta_o shifted by -2 alpha,
    +alpha and +4 alpha; the residual lobe is not shifted
    (disney_hair.cuh:228-254). A_p and N_p do not depend on theta_i, so a
    lobe at tilt alpha and outgoing theta_o is, as a function of theta_i, the
    untilted lobe at theta_o + shift times a constant; and its peak (near
    theta_i = -theta_o') moves against the shift (by about its size for the
    narrow R and TT lobes). A swapped sign or a wrong power of two breaks the proportionality by orders of magnitude."""
    th = np.linspace(-1.3, 1.3, 4001)
    to = math.asin(so)
    for alpha_deg in (3.0, 6.0):
        a = math.radians(alpha_deg)
        for lobe, shift in ((0, -2 * a), (1, a), (2, 4 * a), (3, 0.0)):
            f_t = lobe_curve(so, a, lobe, th)
            f_0 = lobe_curve(math.sin(to + shift), 0.0, lobe, th)
            keep = f_0 > 1e-3 * f_0.max()
            ratio = f_t[keep] / f_0[keep]
            assert np.ptp(ratio) < 1e-4 * ratio.mean(), (lobe, alpha_deg, np.ptp(ratio) / ratio.mean())
            if shift:
                moved = th[np.argmax(f_t)] - th[np.argmax(lobe_curve(so, 0.0, lobe, th))]
                assert moved * -shift > 0, (lobe, moved, shift)
                if lobe < 2:
                    assert abs(moved + shift) < 0.25 * abs(shift), (lobe, moved, shift)


# ---------------------------------------------------------------------------
# Q1: directLighting with an environment and a directional light
# (optix_common.cuh:194, :205-309)

def matched_beta_m(beta_n):
    from scipy.optimize import brentq
    s = 0.626657069 * (0.265 * beta_n + 1.194 * beta_n ** 2 + 5.372 * beta_n ** 22)
    return brentq(lambda b: 4 * (0.726 * b + 0.812 * b * b + 3.7 * b ** 20) ** 2 - s, 0.05, 0.99)


def test_direct_lighting_expectation_q1_q2(orc):
    """One hair hit lit by a procedural environment (importance sampled) and
    one directional light (N = 2 lights). The estimator's mean over 4M random
    streams equals its expectation computed here by quadrature:

      f(w_d) E + int f L w_l d omega
        + (1 / N) [(1 - P_low) int_upper f L w_b d omega + int_lower f L w_b d omega]

    with the power-heuristic weights w_l of the light sample (pdf p_env / N)
    and w_b of the BSDF sample. Q1: the BSDF sample is taken only when the
    environment was picked and is not divided by 1 / N
    (optix_common.cuh:194, :287-293). Q2: a light sample below the fibre
    (probability P_low under p_env) pushes the shared hit point through the
    fibre (:247-250), so a following BSDF sample above it starts inside the
    fibre and is shadowed by it (:179-186). The unbiased weighting (w_l + w_b
    without either quirk) is excluded by far more than the tolerance. beta_m
    is matched so that the BSDF sampler's density equals its pdf (the
    residual lobe is sampled with the logistic scale)."""
    beta_n = 0.5
    beta_m = matched_beta_m(beta_n)
    dl_from, dl_emit = [0.3, 0.8, 0.6], [1.5, 1.2, 0.9]
    sig = [0.2, 0.4, 0.8]
    info = ol.scene_info(has_env=1, env_pdf=1, num_dir_lights=1, dir_from=[dl_from], dir_emit=[dl_emit],
                         beta_m=beta_m, beta_n=beta_n, alpha=ALPHA_2DEG, sigma_a=sig)
    seg = np.array([[-1, 0, 0, 0.05, 1, 0, 0, 0]], np.float32)
    rgba = procedural_env(64, 32)
    osc = ol.OracleScene(info, seg, ol.env_with_tables(rgba))
    # a ray with sin(theta_o) = 0.3 hitting the fibre off-centre
    d = np.array([-0.3, 0.0, math.sqrt(1 - 0.09)])
    o = np.array([0.1, 0.02, 0.0]) - 3 * d
    ray = np.zeros((1, 8), np.float32)
    ray[0, 0:3], ray[0, 4:7] = o, d
    prim, t, u = ol.intersect(seg, ray)
    assert prim[0] == 0
    g = ol.probe_hit(osc, ray, t, prim, u)
    F = np.stack([g["X"][0], g["Y"][0], g["Z"][0]])
    wo_l, h, n_l = g["wo_local"][0], g["h"][0], F @ g["n"][0]
    assert 0.2 < abs(h) < 0.6

    n = 1 << 22
    seeds = np.random.default_rng(21).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    mean, msq = ol.probe_direct(osc, ray, float(t[0]), int(prim[0]), float(u[0]), seeds)
    se = np.sqrt(np.maximum(msq - mean ** 2, 0) / n)

    # quadrature over the sphere in the hair frame: x = sin(theta), d omega = dx dphi
    nx, nph = 2048, 2048
    x = (np.arange(nx) + 0.5) / nx * 2 - 1
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi - np.pi
    X, P = np.meshgrid(x, ph, indexing="ij")
    c = np.sqrt(1 - X * X)
    wl = np.stack([X, c * np.cos(P), c * np.sin(P)], -1).reshape(-1, 3)
    dw = (2 / nx) * (2 * np.pi / nph)
    q = np.zeros((wl.shape[0], 10), np.float32)
    q[:, 0:3], q[:, 3:6], q[:, 6], q[:, 7:10] = wo_l, wl, h, sig
    params = ol.hair_params(beta_m, beta_n, ALPHA_2DEG)
    f, pb = ol.hair_eval(q, params, prec=0)
    L, pe = ol.probe_env(osc, wl @ F)  # to world: rows of F are the frame axes
    N = 2
    pl = pe / N
    w_l = pl ** 2 / (pl ** 2 + pb ** 2)
    w_b = pb ** 2 / (pb ** 2 + pl ** 2)
    fL = f * L
    low = wl @ n_l < 0
    p_low = np.sum(pe[low]) * dw
    light = np.sum(fL * w_l[:, None], 0) * dw
    fLb = fL * w_b[:, None]
    bsdf = ((1 - p_low) * np.sum(fLb[~low], 0) + np.sum(fLb[low], 0)) * dw / N
    wd = np.asarray(dl_from) / np.linalg.norm(dl_from)
    qd = np.zeros((1, 10), np.float32)
    qd[0, 0:3], qd[0, 3:6], qd[0, 6], qd[0, 7:10] = wo_l, F @ wd, h, sig
    fd, _ = ol.hair_eval(qd, params, prec=0)
    direct = fd[0] * np.asarray(dl_emit)
    expect = direct + light + bsdf
    unbiased = direct + np.sum(fL, 0) * dw
    no_q2 = direct + light + np.sum(fLb, 0) * dw / N
    print(f"MC {mean} +- {se}; expected {expect}; without Q2 {no_q2}; unbiased {unbiased}; P_low {p_low:.3f}")
    tol = 4 * se + 1e-3 * expect
    assert np.all(np.abs(mean - expect) < tol)
    # the alternatives differ by far more than the tolerance
    assert np.all(np.abs(unbiased - expect) > 10 * tol)
    assert np.all(np.abs(no_q2 - expect) > 3 * tol)
