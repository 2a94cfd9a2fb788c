"""Pins the FP64 oracle adjoint (path replay with finite-difference BSDF
derivatives) against finite differences of the oracle's own renders.

With directional lights only and direct lighting at the first hit, the image
is a smooth, sampling-free function of (beta_m, beta_n, alpha, sigma_a): the
adjoint must equal central differences of the render to FD accuracy. With
the environment light, MIS and bounces, the sampled directions depend on the
parameters; the oracle's frozen-sampling render (directions, sampling pdfs
and roulette at the base parameters, f and MIS weights at the perturbed
ones) is then the function whose derivative the detached-sampling adjoint
estimates path by path, and finite differences of it must match exactly.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as ol

SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def make_scene(w, h, path_v2, env: bool, strands=600, seed=3):
    import hairmsnn_amd as hm
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["integrator"].update({"width": w, "height": h, "path_v2": path_v2})
    cfg["hair"]["procedural"]["strands"] = strands
    if not env:
        cfg["lights"].pop("environment", None)
    info = hm.parse_scene(json.dumps(cfg))
    seg, envmap = hm.generate_host(cfg)
    e = {"rgba": envmap, **ol.env_tables(envmap, 0)} if envmap is not None and info["has_env"] else None
    ns = int(seg[:, 7].max()) + 1
    rng = np.random.default_rng(seed)
    sigma = (np.asarray(info["sigma_a"], np.float32)[None, :] *
             rng.uniform(0.5, 1.5, (ns, 3)).astype(np.float32))
    return ol.OracleScene(info, seg, e, strand_sigma=sigma), sigma, ns


def loss_weights(w, h, seed=11):
    g = np.random.default_rng(seed).uniform(0.5, 1.5, (h, w, 4)).astype(np.float32)
    g[..., 3] = 0
    return g


def render_loss(osc, wts, spp, per_pixel=False):
    full, _, _ = osc.render(0, 0, osc.s.width, osc.s.height, spp=spp, prec=0)
    v = (full * wts[..., :3]).sum(-1)
    return v if per_pixel else v.sum()


def fd_global(osc, k, h, wts, spp, per_pixel=False):
    name = ("beta_m", "beta_n", "alpha")[k]
    x0 = getattr(osc.s, name)
    xp, xm = np.float32(x0 + h), np.float32(x0 - h)
    setattr(osc.s, name, float(xp))
    lp = render_loss(osc, wts, spp, per_pixel)
    setattr(osc.s, name, float(xm))
    lm = render_loss(osc, wts, spp, per_pixel)
    setattr(osc.s, name, x0)
    return (lp - lm) / float(np.float64(xp) - np.float64(xm))


def fd_sigma(osc, sigma, strands, ch, h, wts, spp):
    base = sigma.copy()
    plus, minus = base.copy(), base.copy()
    plus[strands, ch] += h
    minus[strands, ch] -= h
    step = (plus[strands, ch].astype(np.float64) - minus[strands, ch].astype(np.float64)).mean()
    osc.strand_sigma[:] = plus
    lp = render_loss(osc, wts, spp)
    osc.strand_sigma[:] = minus
    lm = render_loss(osc, wts, spp)
    osc.strand_sigma[:] = base
    return (lp - lm) / step


def test_adjoint_equals_fd_direct_lighting_only():
    w = h = 32  # the pixel count must be a multiple of 128 (scene.cpp image-size check)
    osc, sigma, ns = make_scene(w, h, path_v2=1, env=False)
    wts = loss_weights(w, h)
    spp = 2
    gs, gg, gp = osc.adjoint(wts, spp, ns)
    assert np.abs(gg).max() > 0 and np.isfinite(gs).all()
    np.testing.assert_allclose(gp.sum((0, 1)), gg, rtol=1e-9)
    # central differences of the render have an O(h^2) error (2e-4 relative
    # for beta_m at h = 2e-3); Richardson extrapolation over h and h/2
    # cancels it, leaving FP64 rounding
    for k in range(3):
        f1, f2 = fd_global(osc, k, 1e-3, wts, spp), fd_global(osc, k, 5e-4, wts, spp)
        fd = (4 * f2 - f1) / 3
        print(f"param {k}: adjoint {gg[k]:.8e} fd {fd:.8e}")
        assert abs(gg[k] - fd) <= 1e-6 * abs(fd) + 1e-9 * np.abs(gg).max()
    # sigma_a of every strand at once, and of the three most influential strands
    for ch in range(3):
        f1 = fd_sigma(osc, sigma, np.arange(ns), ch, 2e-3, wts, spp)
        f2 = fd_sigma(osc, sigma, np.arange(ns), ch, 1e-3, wts, spp)
        fd = (4 * f2 - f1) / 3
        assert abs(gs[:, ch].sum() - fd) <= 1e-5 * abs(fd), (ch, gs[:, ch].sum(), fd)
    top = np.argsort(-np.abs(gs[:, 0]))[:3]
    for s in top:
        f1 = fd_sigma(osc, sigma, np.array([s]), 0, 2e-3, wts, spp)
        f2 = fd_sigma(osc, sigma, np.array([s]), 0, 1e-3, wts, spp)
        fd = (4 * f2 - f1) / 3
        assert abs(gs[s, 0] - fd) <= 1e-4 * abs(fd), (s, gs[s, 0], fd)


def freeze(osc, sigma):
    """Sampling (directions, sampling pdfs, roulette) at the current
    parameters; later parameter changes move only f and the MIS weights."""
    osc.sigma0 = sigma.copy()
    osc.s.frozen = 1
    osc.s.beta_m0, osc.s.beta_n0, osc.s.alpha0 = osc.s.beta_m, osc.s.beta_n, osc.s.alpha
    osc.s.strand_sigma0 = ol.ptr(osc.sigma0)


@pytest.mark.parametrize("path_v2", [2, 40])
def test_adjoint_equals_fd_of_frozen_sampling_render(path_v2):
    """Environment + directional light with MIS, bounces and roulette. With
    the sampling frozen at the base parameters the render is a smooth function
    of the parameters whose derivative is exactly what the detached-sampling
    adjoint estimates, path by path: the light-sample and BSDF-sample MIS
    weights through the hair pdf, and every bounce's f."""
    w = h = 32
    spp = 2
    osc, sigma, ns = make_scene(w, h, path_v2=path_v2, env=True)
    wts = loss_weights(w, h)
    gs, gg, gp = osc.adjoint(wts, spp, ns)
    freeze(osc, sigma)
    # M_p switches to its large-argument form at a = 12 (disney_hair.cuh:77-88),
    # a ~1e-3 relative step that a beta_m difference can straddle: compare per
    # pixel and allow a few such pixels
    for k in range(3):
        f1 = fd_global(osc, k, 5e-4, wts, spp, per_pixel=True)
        f2 = fd_global(osc, k, 2.5e-4, wts, spp, per_pixel=True)
        fd = (4 * f2 - f1) / 3
        ok = np.abs(fd - gp[..., k]) <= 1e-5 * np.abs(fd) + 1e-7 * np.abs(gp[..., k]).max()
        print(f"depth {path_v2} param {k}: adjoint {gg[k]:.8e} fd {fd.sum():.8e} pixels agreeing {ok.mean():.4f}")
        assert ok.mean() >= 0.99
        assert abs(gg[k] - fd[ok].sum() - gp[..., k][~ok].sum()) <= 1e-5 * np.abs(gg).max()
    for ch in range(3):
        f1 = fd_sigma(osc, sigma, np.arange(ns), ch, 2e-3, wts, spp)
        f2 = fd_sigma(osc, sigma, np.arange(ns), ch, 1e-3, wts, spp)
        fd = (4 * f2 - f1) / 3
        print(f"depth {path_v2} sigma {ch}: adjoint {gs[:, ch].sum():.8e} fd {fd:.8e}")
        assert abs(gs[:, ch].sum() - fd) <= 1e-5 * abs(fd), (ch, gs[:, ch].sum(), fd)
    # 16 random per-strand parameters (h = 1e-3 and 5e-4, extrapolated)
    rng = np.random.default_rng(7)
    seen = np.flatnonzero(np.abs(gs).max(1) > 0)
    for _ in range(16):
        st, ch = int(rng.choice(seen)), int(rng.integers(3))
        f1 = fd_sigma(osc, sigma, np.array([st]), ch, 1e-3, wts, spp)
        f2 = fd_sigma(osc, sigma, np.array([st]), ch, 5e-4, wts, spp)
        fd = (4 * f2 - f1) / 3
        assert abs(gs[st, ch] - fd) <= 1e-4 * abs(fd) + 1e-9 * np.abs(gs).max(), (st, ch, gs[st, ch], fd)


def test_frozen_render_at_base_parameters_is_the_render():
    osc, sigma, ns = make_scene(32, 32, path_v2=40, env=True)
    a, _, _ = osc.render(0, 0, 32, 32, spp=2, prec=0)
    freeze(osc, sigma)
    b, _, _ = osc.render(0, 0, 32, 32, spp=2, prec=0)
    np.testing.assert_array_equal(a, b)
