"""Seeded synthetic inputs shared by the tests."""
from __future__ import annotations

import numpy as np

ALPHA_2DEG = float(np.float32(np.float32(3.14159) * np.float32(2.0)) / np.float32(180.0))


def unit_sphere(rng, n):
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def bsdf_queries(n: int, seed: int, sigma=None, renderer_like: bool = True) -> np.ndarray:
    """n x 10 float32 queries: wo(3), wi(3), h, sigma_a(3).

    renderer_like: wo has y = 0 and z >= 0, as the hair frame makes it."""
    rng = np.random.default_rng(seed)
    if renderer_like:
        so = rng.uniform(-0.995, 0.995, n)
        wo = np.stack([so, np.zeros(n), np.sqrt(1 - so * so)], 1)
    else:
        wo = unit_sphere(rng, n)
    wi = unit_sphere(rng, n)
    h = rng.uniform(-0.999, 0.999, n)
    if sigma is None:
        sig = rng.uniform(0.01, 1.0, (n, 3))
    else:
        sig = np.broadcast_to(np.asarray(sigma, np.float64), (n, 3))
    q = np.concatenate([wo, wi, h[:, None], sig], 1).astype(np.float32)
    # renormalise in FP32 so the stored directions are unit length in FP32
    for a, b in ((0, 3), (3, 6)):
        v = q[:, a:b].astype(np.float64)
        q[:, a:b] = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    return q


def uniforms(n: int, k: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).random((n, k), dtype=np.float32)
