"""Oracle RNG against the closed-form integer arithmetic of lcg_random.cuh:15-66,
restated here independently in Python."""
import numpy as np


def rotl(v, r):
    return ((v << r) | (v >> (32 - r))) & 0xFFFFFFFF


def mix(h, k):
    k = (k * 0xCC9E2D51) & 0xFFFFFFFF
    k = rotl(k, 15)
    k = (k * 0x1B873593) & 0xFFFFFFFF
    h ^= k
    return (rotl(h, 13) * 5 + 0xE6546B64) & 0xFFFFFFFF


def fmix(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    h ^= h >> 16
    return h


def test_murmur3_known_answer(orc):
    # Murmur3_x86_32 of the empty key with seed 0 is 0; the finaliser is a
    # bijection with fmix(0) == 0 (published reference values).
    assert fmix(0) == 0
    # murmur3_32("abcd" as one block, seed 0) = 0x43ED676A; with the length
    # mixed in (h ^= 4) before finalisation.
    k = int.from_bytes(b"abcd", "little")
    assert fmix(mix(0, k) ^ 4) == 0x43ED676A


def test_seed_matches_closed_form(orc):
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y = int(rng.integers(0, 2048)), int(rng.integers(0, 2048))
        w = int(rng.integers(1, 4096))
        frame = int(rng.integers(0, 100000)) + 10007
        expect = fmix(mix(mix(0, (x + y * w) & 0xFFFFFFFF), frame))
        assert orc.orc_path_seed(frame, x, y, w) == expect


def test_lcg_stream_and_uniforms(orc):
    from tests.oracle_lib import ptr
    seed = 0xDEADBEEF
    n = 1000
    u32 = np.empty(n, np.uint32)
    f32 = np.empty(n, np.float32)
    orc.orc_lcg_stream(seed, n, ptr(u32), ptr(f32))
    s = seed
    for i in range(n):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        assert int(u32[i]) == s
    # float(u32) * 2^-32 with round-to-nearest conversion
    np.testing.assert_array_equal(f32, (u32.astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float32))
    assert np.all(f32 >= 0) and np.all(f32 <= 1)


def test_uniform_can_be_exactly_one(orc):
    """u32 >= 2^32 - 128 rounds to 2^32 in FP32, so the draw is exactly 1.0."""
    from tests.oracle_lib import ptr
    # choose a state whose successor is 0xFFFFFFF0
    target = 0xFFFFFFF0
    inv = pow(1664525, -1, 2 ** 32)
    prev = ((target - 1013904223) * inv) & 0xFFFFFFFF
    f = np.empty(1, np.float32)
    orc.orc_lcg_stream(prev, 1, None, ptr(f))
    assert f[0] == 1.0
