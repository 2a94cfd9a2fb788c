"""Path-replay adjoint on the GPU (k_adjoint) against the FP64 oracle adjoint,
against central differences of GPU renders, and through the torch.autograd
harness."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu
SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def scene_cfg(w, h, path_v2, env=True, strands=600):
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["integrator"].update({"width": w, "height": h, "path_v2": path_v2})
    cfg["hair"]["procedural"]["strands"] = strands
    if not env:
        cfg["lights"].pop("environment", None)
    return cfg


def strand_sigma(info, ns, seed=3):
    rng = np.random.default_rng(seed)
    return (np.asarray(info["sigma_a"], np.float32)[None, :] *
            rng.uniform(0.5, 1.5, (ns, 3)).astype(np.float32))


def loss_weights(w, h, seed=11):
    g = np.random.default_rng(seed).uniform(0.5, 1.5, (h, w, 4)).astype(np.float32)
    g[..., 3] = 0
    return g


@pytest.mark.parametrize("path_v2,need", [(1, 0.97), (2, 0.88)])
def test_matches_fp64_oracle_adjoint_per_pixel(gpu, path_v2, need):
    """Same paths, same seeds: each pixel's d/d(beta_m, beta_n, alpha) must
    match the FP64 oracle (dual numbers vs finite differences) wherever the
    FP32 and FP64 paths have not decorrelated (the per-pixel agreement of
    the renders themselves is about 0.99 / 0.95 at these depths)."""
    w = h = 64
    spp = 2
    with gpu.Scene(json_text=scene_cfg(w, h, path_v2)) as sc:
        ns = sc.num_strands
        sig = strand_sigma(sc.info, ns)
        sc.set_strand_sigma_a(sig)
        wts = loss_weights(w, h)
        gs, gg, gp = sc.adjoint(wts, spp, per_pixel=True)
        osc = ol.OracleScene(sc.info, sc.segments(), sc.env(), strand_sigma=sig)
    ogs, ogg, ogp = osc.adjoint(wts, spp, ns)
    scale = np.abs(ogp).max(axis=(0, 1))
    close = np.abs(gp - ogp) <= 2e-3 * np.abs(ogp) + 1e-4 * scale
    frac = close.all(-1).mean()
    print(f"depth {path_v2}: GPU {gg} oracle {ogg}; pixels agreeing {frac:.4f}")
    assert frac >= need
    np.testing.assert_allclose(gp.sum((0, 1)), gg, rtol=1e-4, atol=1e-6 * np.abs(gg).max())
    # per-strand sigma_a: strands seen only through agreeing pixels match closely
    ssc = np.abs(ogs).max()
    sclose = (np.abs(gs - ogs) <= 5e-3 * np.abs(ogs) + 1e-4 * ssc).all(-1).mean()
    corr = np.corrcoef(gs.ravel(), ogs.ravel())[0, 1]
    print(f"strands agreeing {sclose:.4f}, correlation {corr:.6f}")
    assert sclose >= need - 0.05 and corr > 0.99


def render_loss(sc, wts, spp, per_pixel=False):
    img = sc.render_path_tracing(spp)[..., :3].astype(np.float64)
    v = (img * wts[..., :3]).sum(-1)
    return v if per_pixel else v.sum()


def fd_param(sc, k, step, wts, spp, per_pixel=False):
    p = [float(sc.info["beta_m"]), float(sc.info["beta_n"]), float(sc.info["alpha"])]
    base = list(p)
    xp, xm = np.float32(p[k] + step), np.float32(p[k] - step)
    p[k] = float(xp)
    sc.set_hair_params(*p)
    lp = render_loss(sc, wts, spp, per_pixel)
    p[k] = float(xm)
    sc.set_hair_params(*p)
    lm = render_loss(sc, wts, spp, per_pixel)
    sc.set_hair_params(*base)
    return (lp - lm) / (float(xp) - float(xm))


def test_matches_central_differences_direct_lighting(gpu):
    """Directional light, direct lighting at the first hit: the FP32 render is
    sampling-free in the parameters, so Richardson-extrapolated central
    differences must match the adjoint per pixel up to FP32 rounding."""
    w = h = 64
    spp = 2
    with gpu.Scene(json_text=scene_cfg(w, h, 1, env=False)) as sc:
        ns = sc.num_strands
        sig = strand_sigma(sc.info, ns)
        sc.set_strand_sigma_a(sig)
        wts = loss_weights(w, h)
        gs, gg, gp = sc.adjoint(wts, spp, per_pixel=True)
        for k in range(3):
            f1 = fd_param(sc, k, 4e-3, wts, spp, per_pixel=True)
            f2 = fd_param(sc, k, 2e-3, wts, spp, per_pixel=True)
            fd = (4 * f2 - f1) / 3
            ok = np.abs(fd - gp[..., k]) <= 5e-3 * np.abs(gp[..., k]) + 2e-4 * np.abs(gp[..., k]).max()
            # a few pixels straddle M_p's switch to its large-argument form
            # (a ~1e-3 step in f) within the difference step: they are the
            # disagreeing ones, and the rest must add up
            print(f"param {k}: adjoint {gg[k]:.6e} fd {fd.sum():.6e} pixels agreeing {ok.mean():.4f}")
            assert ok.mean() >= 0.97
            assert abs(fd[ok].sum() + gp[..., k][~ok].sum() - gg[k]) <= 2e-3 * np.abs(gg).max()
        for ch in range(3):
            lp, lm = sig.copy(), sig.copy()
            lp[:, ch] += 2e-3
            lm[:, ch] -= 2e-3
            sc.set_strand_sigma_a(lp)
            a = render_loss(sc, wts, spp)
            sc.set_strand_sigma_a(lm)
            b = render_loss(sc, wts, spp)
            step = float((lp[:, ch].astype(np.float64) - lm[:, ch]).mean())
            fd = (a - b) / step
            print(f"sigma {ch}: adjoint {gs[:, ch].sum():.6e} fd {fd:.6e}")
            assert abs(fd - gs[:, ch].sum()) <= 3e-3 * abs(fd)
        # 64 random per-strand parameters among the strands the image sees,
        # central differences with h = 1e-3
        rng = np.random.default_rng(5)
        seen = np.flatnonzero(np.abs(gs).max(1) > 1e-3 * np.abs(gs).max())
        picks = [(int(rng.choice(seen)), int(rng.integers(3))) for _ in range(64)]
        bad = 0
        for st, ch in picks:
            lp, lm = sig.copy(), sig.copy()
            lp[st, ch] += 1e-3
            lm[st, ch] -= 1e-3
            sc.set_strand_sigma_a(lp)
            a = render_loss(sc, wts, spp)
            sc.set_strand_sigma_a(lm)
            b = render_loss(sc, wts, spp)
            fd = (a - b) / (float(lp[st, ch]) - float(lm[st, ch]))
            if abs(fd - gs[st, ch]) > 1e-2 * abs(gs[st, ch]) + 1e-4 * np.abs(gs).max():
                bad += 1
                print(f"strand {st} ch {ch}: adjoint {gs[st, ch]:.6e} fd {fd:.6e}")
        assert bad <= 2
        sc.set_strand_sigma_a(sig)


def test_unbiased_against_crn_differences_full_depth(gpu):
    """Full-depth paths with environment + directional light and MIS: per-pixel
    central differences with common random numbers and the detached-sampling
    adjoint estimate the same derivative of the expected image; their summed
    difference must stay within 4 standard errors (plus the O(h^2) bias of the
    difference step)."""
    w = h = 64
    spp = 1024
    with gpu.Scene(json_text=scene_cfg(w, h, 40)) as sc:
        wts = loss_weights(w, h)
        gs, gg, gp = sc.adjoint(wts, spp, per_pixel=True)
        for k in range(3):
            fd = fd_param(sc, k, 5e-3, wts, spp, per_pixel=True)
            d = (fd - gp[..., k]).ravel()
            se = d.std(ddof=1) * np.sqrt(d.size)
            print(f"param {k}: adjoint {gg[k]:.4f} crn-fd {fd.sum():.4f} diff {d.sum():.4f} se {se:.4f}")
            assert abs(d.sum()) <= 4 * se + 0.01 * abs(gg[k])


def test_torch_autograd_harness(gpu):
    import torch
    from hairmsnn_amd.diff import render_hair
    w = h = 32
    spp = 2
    with gpu.Scene(json_text=scene_cfg(w, h, 3)) as sc:
        ns = sc.num_strands
        sig = torch.tensor(strand_sigma(sc.info, ns), dtype=torch.float64, requires_grad=True)
        prm = torch.tensor([float(sc.info["beta_m"]), float(sc.info["beta_n"]), float(sc.info["alpha"])],
                           dtype=torch.float64, requires_grad=True)
        wts = torch.from_numpy(loss_weights(w, h)[..., :3].astype(np.float64))
        img = render_hair(sc, sig, prm, spp)
        assert img.shape == (h, w, 3) and img.dtype == torch.float64
        loss = (img * wts).sum()
        loss.backward()
        gs, gg = sc.adjoint(wts.numpy().astype(np.float32), spp)
    np.testing.assert_allclose(prm.grad.numpy(), gg, rtol=1e-6)
    np.testing.assert_allclose(sig.grad.numpy(), gs, rtol=1e-6, atol=1e-12)
    assert np.isfinite(gs).all() and np.abs(gg).max() > 0


def c1_cfg(w, h, path_v2, env=True):
    cfg = json.loads((SCENES / "c1_curly.json").read_text())
    cfg["integrator"].update({"width": w, "height": h, "path_v2": path_v2})
    if not env:
        cfg["lights"].pop("environment", None)
    return cfg


def c3_sigma(ns):
    """The differentiable-appearance configuration's per-strand absorption:
    a uniform (0.01, 0.5) draw per strand and channel, seed 3."""
    return np.random.default_rng(3).uniform(0.01, 0.5, (ns, 3)).astype(np.float32)


def test_c3_on_c1_geometry_matches_fp64_oracle(gpu):
    """C1's geometry (50,000 strands x 64 segments) with 50,000 x 3 per-strand
    sigma_a, the procedural environment and the directional light, direct
    lighting: per-pixel d/d(beta_m, beta_n, alpha) and per-strand d/d sigma_a
    against the FP64 oracle adjoint on the same paths."""
    w = h = 64
    spp = 2
    with gpu.Scene(json_text=c1_cfg(w, h, 1)) as sc:
        ns = sc.num_strands
        assert ns == 50_000
        sig = c3_sigma(ns)
        sc.set_strand_sigma_a(sig)
        wts = loss_weights(w, h)
        gs, gg, gp = sc.adjoint(wts, spp, per_pixel=True)
        osc = ol.OracleScene(sc.info, sc.segments(), sc.env(), strand_sigma=sig, mesh=sc.mesh())
    ogs, ogg, ogp = osc.adjoint(wts, spp, ns)
    scale = np.abs(ogp).max(axis=(0, 1))
    close = np.abs(gp - ogp) <= 2e-3 * np.abs(ogp) + 1e-4 * scale
    frac = close.all(-1).mean()
    print(f"C1 direct: GPU {gg} oracle {ogg}; pixels agreeing {frac:.4f}")
    assert frac >= 0.95
    seen = np.abs(ogs).max(1) > 0
    ssc = np.abs(ogs).max()
    sclose = (np.abs(gs - ogs) <= 5e-3 * np.abs(ogs) + 1e-4 * ssc).all(-1)[seen].mean()
    corr = np.corrcoef(gs[seen].ravel(), ogs[seen].ravel())[0, 1]
    print(f"strands seen {seen.sum()}, agreeing {sclose:.4f}, correlation {corr:.6f}")
    assert seen.sum() > 500 and sclose >= 0.9 and corr > 0.99


def test_c3_on_c1_geometry_central_differences(gpu):
    """C1 geometry lit by the directional light only (the render is then
    sampling-free in the parameters at the first vertex): 64 random
    per-strand sigma_a entries among the strands the image sees, central
    differences with h = 1e-3 against the adjoint."""
    w = h = 64
    spp = 2
    with gpu.Scene(json_text=c1_cfg(w, h, 1, env=False)) as sc:
        ns = sc.num_strands
        sig = c3_sigma(ns)
        sc.set_strand_sigma_a(sig)
        wts = loss_weights(w, h)
        gs, gg, _ = sc.adjoint(wts, spp, per_pixel=True)
        rng = np.random.default_rng(7)
        seen = np.flatnonzero(np.abs(gs).max(1) > 1e-3 * np.abs(gs).max())
        assert seen.size > 30
        picks = [(int(rng.choice(seen)), int(rng.integers(3))) for _ in range(64)]
        bad = 0
        for st, ch in picks:
            lp, lm = sig.copy(), sig.copy()
            lp[st, ch] += 1e-3
            lm[st, ch] -= 1e-3
            sc.set_strand_sigma_a(lp)
            a = render_loss(sc, wts, spp)
            sc.set_strand_sigma_a(lm)
            b = render_loss(sc, wts, spp)
            fd = (a - b) / (float(lp[st, ch]) - float(lm[st, ch]))
            if abs(fd - gs[st, ch]) > 1e-2 * abs(gs[st, ch]) + 1e-4 * np.abs(gs).max():
                bad += 1
                print(f"strand {st} ch {ch}: adjoint {gs[st, ch]:.6e} fd {fd:.6e}")
        sc.set_strand_sigma_a(sig)
    print(f"64 central differences: {64 - bad} agree")
    assert bad <= 2
