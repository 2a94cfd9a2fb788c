"""hmsnn_mlp_train_step_dev_noalloc (a training step whose scratch is a
static device array) against hmsnn_mlp_train_step_dev (the same kernels with
buffers allocated on demand): from the same weights, two steps in a row, the
second on a smaller batch, give the same weights bit for bit and the same
loss to rounding; batches beyond the scratch's 16384 records are refused."""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_lib as ol
from tests.test_oracle_mlp import records, weights

pytestmark = pytest.mark.gpu


def test_noalloc_train_step_matches_allocating_step(gpu):
    lib = gpu._lib.load()
    hs = []
    for prec in (0, 0):
        h = C.c_void_p()
        gpu._lib.check(lib.hmsnn_mlp_create(prec, 1, C.byref(h)))
        hs.append(h)
    a, b = hs
    w = weights(61)
    for h in hs:
        gpu._lib.check(lib.hmsnn_mlp_set_weights(h, C.c_void_p(w.ctypes.data), w.size))
    loss_dev = gpu.DeviceBuffer(4)
    try:
        for step, n in enumerate((16384, 1000)):
            r = records(n, 70 + step)
            t = np.random.default_rng(80 + step).uniform(0.05, 2.0, (n, 3)).astype(np.float32)
            di = gpu.DeviceBuffer.from_numpy(r)
            dt = gpu.DeviceBuffer.from_numpy(t)
            la = C.c_float()
            gpu._lib.check(lib.hmsnn_mlp_train_step_dev(a, di.ptr, dt.ptr, n, 0, C.c_float(1e-2), C.byref(la), None))
            gpu._lib.check(lib.hmsnn_mlp_train_step_dev_noalloc(b, di.ptr, dt.ptr, n, 0, C.c_float(1e-2),
                                                                loss_dev.ptr, None))
            gpu._lib.check(lib.hmsnn_dev_synchronize())
            lb = float(loss_dev.to_numpy(np.float32, (1,))[0])
            wa = np.empty(ol.MLP_WEIGHTS, np.float32)
            wb = np.empty(ol.MLP_WEIGHTS, np.float32)
            gpu._lib.check(lib.hmsnn_mlp_get_weights(a, C.c_void_p(wa.ctypes.data), wa.size))
            gpu._lib.check(lib.hmsnn_mlp_get_weights(b, C.c_void_p(wb.ctypes.data), wb.size))
            print(f"step {step} (n = {n}): loss {la.value:.6f} / {lb:.6f}, weights moved "
                  f"{np.abs(wa - w).max():.3e}")
            # the same kernels sum the batch loss in a fixed order; the
            # allocating step divides it on the host, this one on the device
            assert lb == pytest.approx(la.value, rel=1e-5)
            np.testing.assert_array_equal(wb, wa)
            assert not np.array_equal(wa, w)
            w = wa
        r = records(16385, 90)
        di = gpu.DeviceBuffer.from_numpy(r)
        dt = gpu.DeviceBuffer.from_numpy(np.ones((16385, 3), np.float32))
        rc = lib.hmsnn_mlp_train_step_dev_noalloc(b, di.ptr, dt.ptr, 16385, 0, C.c_float(1e-2), loss_dev.ptr, None)
        assert rc == gpu._lib.HMSNN_EINVAL and b"16384" in lib.hmsnn_last_error()
    finally:
        for h in hs:
            lib.hmsnn_mlp_destroy(h)
