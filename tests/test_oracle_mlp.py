"""Pins the oracle's radiance-cache network: encoding closed forms, gradient
against central finite differences, FP32 twin against FP64."""
import math

import numpy as np
import pytest

from tests import oracle_lib as ol


def records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 12), np.float32)
    r[:, 0:3] = rng.uniform(-0.5, 0.5, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        r[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    r[:, 9:12] = rng.uniform(0.01, 0.5, (n, 3))
    return r


def weights(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.empty(ol.MLP_WEIGHTS, np.float32)
    lim = math.sqrt(6 / 128) * scale
    w[:] = rng.uniform(-lim, lim, ol.MLP_WEIGHTS)
    return w


def test_encoding_closed_form(orc):
    r = records(64, 1)
    f = ol.mlp_encode(r)
    assert f.shape == (64, 64)
    for c in range(3):
        for l in range(6):
            a = (2 ** l) * math.pi * r[:, c].astype(np.float64)
            np.testing.assert_allclose(f[:, 12 * c + 2 * l], np.sin(a), atol=1e-12)
            np.testing.assert_allclose(f[:, 12 * c + 2 * l + 1], np.cos(a), atol=1e-12)
    wo = r[:, 3:6].astype(np.float64)
    np.testing.assert_allclose(f[:, 36], np.arccos(np.clip(wo[:, 2], -1, 1)) / math.pi, atol=1e-12)
    np.testing.assert_allclose(f[:, 37], np.arctan2(wo[:, 1], wo[:, 0]) / (2 * math.pi) + 0.5, atol=1e-12)
    np.testing.assert_allclose(f[:, 40:43], r[:, 9:12], atol=0)
    assert np.all(f[:, 43] == 1.0) and np.all(f[:, 44:] == 0.0)


def _fd_check(w, r, t, loss, tol, h_rel=1e-4):
    # A probe whose +-h window crosses a ReLU kink sees a one-sided slope; at
    # most a tenth of the probes may miss the tight bound, none a loose one.
    l0, g = ol.mlp_loss_grad(w, r, t, loss=loss)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.integers(k * 4096, (k + 1) * 4096, 6) for k in range(5)] +
                         [rng.integers(5 * 4096, ol.MLP_WEIGHTS, 6)])
    checked, loose = 0, 0
    for i in idx:
        h = h_rel * max(1.0, abs(float(w[i])))
        wp, wm = w.copy(), w.copy()
        wp[i] = np.float32(float(w[i]) + h)
        wm[i] = np.float32(float(w[i]) - h)
        hh = (float(wp[i]) - float(wm[i])) / 2
        lp, _ = ol.mlp_loss_grad(wp, r, t, loss=loss)
        lm, _ = ol.mlp_loss_grad(wm, r, t, loss=loss)
        fd = (lp - lm) / (2 * hh)
        if abs(fd) < 1e-12 and abs(g[i]) < 1e-12:
            continue  # dead ReLU path
        err = abs(fd - g[i])
        assert err <= max(0.05, 10 * tol) * abs(fd) + 1e-10, (i, fd, g[i])
        loose += err > tol * abs(fd) + 1e-10
        checked += 1
    assert checked >= 18 and loose <= checked // 10, (checked, loose)


def test_gradient_central_differences_l2(orc):
    r = records(64, 2)
    w = weights(3)
    t = np.random.default_rng(4).uniform(0.1, 2.0, (64, 3)).astype(np.float32)
    _fd_check(w, r, t, loss=2, tol=2e-5)


@pytest.mark.parametrize("loss", [0, 1])
def test_gradient_relative_losses(orc, loss):
    """Relative L2 treats its denominator as a constant (stop-gradient). With
    every target 1 % below its prediction the moving denominator contributes
    a term ~1 % of the gradient, so central differences agree to a few percent."""
    r = records(64, 2)
    w = weights(3)
    y = ol.mlp_forward(w, r, 0)
    t = (0.99 * y).astype(np.float32)
    _fd_check(w, r, t, loss=loss, tol=0.05)


@pytest.mark.parametrize("loss", [0, 1, 2])
def test_loss_value(orc, loss):
    r = records(256, 9)
    w = weights(10)
    t = np.random.default_rng(11).uniform(0.0, 2.0, (256, 3)).astype(np.float32)
    l, _ = ol.mlp_loss_grad(w, r, t, loss=loss)
    y = ol.mlp_forward(w, r, 0)
    d = y - t
    if loss == 0:
        lum = y @ np.array([0.212671, 0.715160, 0.072169])
        expect = np.mean(d * d / (lum[:, None] ** 2 + 0.01))
    elif loss == 1:
        expect = np.mean(d * d / (y * y + 0.01))
    else:
        expect = np.mean(d * d)
    assert l == pytest.approx(expect, rel=1e-10)


def test_fp32_twin(orc):
    r = records(4096, 6)
    w = weights(7)
    y64 = ol.mlp_forward(w, r, 0)
    y32 = ol.mlp_forward(w, r, 1)
    np.testing.assert_allclose(y32, y64, rtol=1e-5)
