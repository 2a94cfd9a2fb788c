"""OBJ head meshes (loadOBJ, model.cpp:233-358; triangleMeshCH,
optix_common.cuh:527-556): the reader's triangles, normals and material
colours against the written file, and the oracle's ray / triangle test
against the closed-form hit of a ray through a known barycentric point."""
import numpy as np
import pytest

from tests import oracle_lib as ol
from tests.objfile import icosphere, write_obj


def test_reader_returns_the_written_triangles(tmp_path):
    import hairmsnn_amd as hm
    P, N, F = icosphere(1, radius=0.5, center=(0.1, -0.2, 0.3))
    for negative in (False, True):
        p = tmp_path / f"head{int(negative)}.obj"
        write_obj(p, P, N, F, negative=negative, quads=True)
        m = hm.load_obj(p)
        assert m["v"].shape == (len(F) + 2, 3, 3)
        np.testing.assert_allclose(m["v"][:len(F)], P[F], atol=1e-6)
        np.testing.assert_allclose(m["n"][:len(F)], N[F], atol=1e-6)
        half = len(F) // 2
        np.testing.assert_allclose(m["kd"][:half], [[0.4, 0.3, 0.2]] * half, rtol=1e-6)
        np.testing.assert_allclose(m["kd"][half:len(F)], [[0.05, 0.05, 0.05]] * (len(F) - half), rtol=1e-6)
        # the quad a b c a fans into (a, b, c) and (a, c, a)
        a, b, c = F[0]
        np.testing.assert_allclose(m["v"][len(F)], P[[a, b, c]], atol=1e-6)
        np.testing.assert_allclose(m["v"][len(F) + 1], P[[a, c, a]], atol=1e-6)


def test_reader_rejects_bad_files(tmp_path):
    import hairmsnn_amd as hm
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    tex = tmp_path / "tex.obj"
    (tmp_path / "tex.mtl").write_text("newmtl a\nmap_Kd skin.png\n")
    tex.write_text("mtllib tex.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    for f, code in ((bad, 2), (tex, 5), (tmp_path / "none.obj", 2)):
        with pytest.raises(hm.HmsnnError) as e:
            hm.load_obj(f)
        assert e.value.code == code


def test_oracle_triangle_test_hits_known_points():
    """Rays aimed at known barycentric points hit there (t, u, v to FP32
    accuracy); rays aimed just outside an edge or pointing away miss."""
    lib = ol.load_oracle()
    rng = np.random.default_rng(0)
    n = 400
    tri = rng.normal(size=(n, 3, 3)).astype(np.float32)
    bu, bv = rng.uniform(0.05, 0.45, n), rng.uniform(0.05, 0.45, n)
    outside = np.arange(n) % 4 == 3
    bu[outside] = -0.05  # just across the edge opposite vertex 1
    away = np.arange(n) % 4 == 2
    target = (1 - bu - bv)[:, None] * tri[:, 0] + bu[:, None] * tri[:, 1] + bv[:, None] * tri[:, 2]
    orig = target + rng.normal(size=(n, 3)) * 2
    d = (target - orig) / np.linalg.norm(target - orig, axis=1, keepdims=True)
    d[away] *= -1
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = orig, d
    hit = np.zeros(n, np.int32)
    t, u, v = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    lib.orc_tri_intersect(n, ol.ptr(np.ascontiguousarray(tri)), ol.ptr(rays), ol.ptr(hit), ol.ptr(t), ol.ptr(u),
                          ol.ptr(v))
    inside = ~outside & ~away
    assert hit[inside].all() and not hit[outside | away].any()
    np.testing.assert_allclose(u[inside], bu[inside], atol=2e-5)
    np.testing.assert_allclose(v[inside], bv[inside], atol=2e-5)
    dist = np.linalg.norm(target - orig, axis=1)
    np.testing.assert_allclose(t[inside], dist[inside], rtol=1e-5)
