"""MFMA radiance-cache MLP against the oracle network (FP32 / FP64 CPU):
encoding, forward (FP16 <= 1e-2, BF16 <= 2e-2 relative), weight gradient,
and one Adam step."""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_lib as ol
from tests.test_oracle_mlp import records, weights

pytestmark = pytest.mark.gpu


class Mlp:
    def __init__(self, hm, precision=0, seed=1):
        self.hm = hm
        self.lib = hm._lib.load()
        h = C.c_void_p()
        hm._lib.check(self.lib.hmsnn_mlp_create(precision, seed, C.byref(h)))
        self.h = h

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self.hm._lib.check(self.lib.hmsnn_mlp_set_weights(self.h, C.c_void_p(w.ctypes.data), w.size))

    def get_weights(self):
        w = np.empty(ol.MLP_WEIGHTS, np.float32)
        self.hm._lib.check(self.lib.hmsnn_mlp_get_weights(self.h, C.c_void_p(w.ctypes.data), w.size))
        return w

    def forward(self, r):
        n = r.shape[0]
        di = self.hm.DeviceBuffer.from_numpy(r.astype(np.float32))
        do = self.hm.DeviceBuffer(12 * n)
        self.hm._lib.check(self.lib.hmsnn_mlp_forward_dev(self.h, di.ptr, n, do.ptr, None))
        self.hm._lib.check(self.lib.hmsnn_dev_synchronize())
        return do.to_numpy(np.float32, (n, 3))

    def grad(self, r, t, loss):
        n = r.shape[0]
        di = self.hm.DeviceBuffer.from_numpy(r.astype(np.float32))
        dt = self.hm.DeviceBuffer.from_numpy(t.astype(np.float32))
        dg = self.hm.DeviceBuffer(4 * ol.MLP_WEIGHTS)
        lo = C.c_float()
        self.hm._lib.check(self.lib.hmsnn_mlp_backward_dev(self.h, di.ptr, dt.ptr, n, loss, dg.ptr, C.byref(lo), None))
        self.hm._lib.check(self.lib.hmsnn_dev_synchronize())
        return lo.value, dg.to_numpy(np.float32, (ol.MLP_WEIGHTS,))

    def step(self, r, t, loss, lr):
        n = r.shape[0]
        di = self.hm.DeviceBuffer.from_numpy(r.astype(np.float32))
        dt = self.hm.DeviceBuffer.from_numpy(t.astype(np.float32))
        lo = C.c_float()
        self.hm._lib.check(self.lib.hmsnn_mlp_train_step_dev(self.h, di.ptr, dt.ptr, n, loss, lr, C.byref(lo), None))
        self.hm._lib.check(self.lib.hmsnn_dev_synchronize())
        return lo.value

    def __del__(self):
        try:
            self.lib.hmsnn_mlp_destroy(self.h)
        except Exception:
            pass


def test_encoding_matches_oracle(gpu):
    r = records(4096, 21)
    di = gpu.DeviceBuffer.from_numpy(r)
    df = gpu.DeviceBuffer(4 * 64 * r.shape[0])
    lib = gpu._lib.load()
    gpu._lib.check(lib.hmsnn_mlp_encode_dev(di.ptr, r.shape[0], df.ptr, None))
    gpu._lib.check(lib.hmsnn_dev_synchronize())
    f = df.to_numpy(np.float32, (r.shape[0], 64))
    np.testing.assert_allclose(f, ol.mlp_encode(r), atol=3e-5)


@pytest.mark.parametrize("precision,tol", [(0, 1e-2), (1, 2e-2)])
def test_forward_matches_oracle(gpu, precision, tol):
    r = records(1 << 16, 22)
    w = weights(23)
    m = Mlp(gpu, precision)
    m.set_weights(w)
    np.testing.assert_array_equal(m.get_weights(), w)
    y = m.forward(r).astype(np.float64)
    y32 = ol.mlp_forward(w, r, 1)
    y64 = ol.mlp_forward(w, r, 0)
    e32 = np.abs(y - y32) / np.abs(y32)
    e64 = np.abs(y - y64) / np.abs(y64)
    print(f"precision {precision}: vs FP32 p99 {np.quantile(e32, 0.99):.2e} max {e32.max():.2e}; "
          f"vs FP64 p99 {np.quantile(e64, 0.99):.2e}")
    assert np.quantile(e32, 0.99) <= tol
    assert np.quantile(e64, 0.99) <= tol
    assert np.all(np.isfinite(y))


def test_fp8_inference_against_fp64_and_fp16(gpu):
    """FP8-E4M3 inference (per-layer power-of-two scales, BF16 calibration)
    on 2^20 seeded inputs against the FP64 oracle, beside FP16 on the same
    weights."""
    n = 1 << 20
    r = records(n, 31)
    w = weights(32)
    y64 = ol.mlp_forward(w, r, 0)
    out = {}
    for prec in (0, 2):
        m = Mlp(gpu, prec)
        m.set_weights(w)
        out[prec] = m.forward(r).astype(np.float64)
        bad = ~np.isfinite(out[prec]).all(-1)
        if bad.any():
            i = np.flatnonzero(bad)[:4]
            print(f"precision {prec}: {bad.sum()} non-finite outputs, e.g. rows {i}: {out[prec][i]} "
                  f"(FP64 {y64[i]}) inputs {r[i]}")
    assert all(np.isfinite(out[p]).all() for p in out)
    e8 = np.abs(out[2] - y64) / np.abs(y64)
    e16 = np.abs(out[0] - y64) / np.abs(y64)
    e8_16 = np.abs(out[2] - out[0]) / np.abs(out[0])
    print(f"FP8 vs FP64: median {np.median(e8):.2e} p99 {np.quantile(e8, 0.99):.2e} max {e8.max():.2e}; "
          f"FP16 vs FP64: median {np.median(e16):.2e} p99 {np.quantile(e16, 0.99):.2e}; "
          f"FP8 vs FP16 p99 {np.quantile(e8_16, 0.99):.2e}")
    assert np.median(e8) <= 2e-2
    assert np.quantile(e8, 0.99) <= 1e-1


@pytest.mark.parametrize("loss", [0, 2])
def test_gradient_matches_oracle(gpu, loss):
    n = 1 << 14
    r = records(n, 24)
    w = weights(25)
    t = np.random.default_rng(26).uniform(0.2, 3.0, (n, 3)).astype(np.float32)
    m = Mlp(gpu, 0)
    m.set_weights(w)
    lg, g = m.grad(r, t, loss)
    lo, go = ol.mlp_loss_grad(w, r, t, loss=loss, prec=0)
    assert lg == pytest.approx(lo, rel=1e-2)
    for l in range(6):
        sl = slice(l * 4096, (l + 1) * 4096) if l < 5 else slice(5 * 4096, ol.MLP_WEIGHTS)
        rel = np.linalg.norm(g[sl] - go[sl]) / np.linalg.norm(go[sl])
        print(f"loss {loss} layer {l}: relative gradient error {rel:.2e}")
        assert rel <= 2e-2


def test_adam_step_and_training_reduces_loss(gpu):
    n = 1 << 14
    r = records(n, 27)
    w = weights(28)
    # a smooth target the network can fit
    t = (0.5 + 0.4 * np.sin(3 * r[:, 0:3])).astype(np.float32)
    m = Mlp(gpu, 0)
    m.set_weights(w)
    _, g = m.grad(r, t, 2)
    lr = 1e-3
    l0 = m.step(r, t, 2, lr)
    w1 = m.get_weights().astype(np.float64)
    gp = g.astype(np.float64) + 1e-6 * w
    # first Adam step with bias correction: w -= lr * g / (|g| + eps)
    expect = w - lr * gp / (np.abs(gp) + 1e-15)
    big = np.abs(gp) > 1e-6
    np.testing.assert_allclose(w1[big], expect[big], atol=2e-6)
    losses = [l0]
    for _ in range(200):
        losses.append(m.step(r, t, 2, 5e-3))
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.2 * losses[0]
