"""OBJ test meshes: a subdivided icosphere written as Wavefront OBJ with
vertex normals, texture coordinates and a material library."""
import numpy as np


def icosphere(subdiv=2, radius=1.0, center=(0.0, 0.0, 0.0)):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]
        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    P = np.array(v) * radius + np.asarray(center)
    N = np.array(v)
    return P, N, np.array(f)


def write_obj(path, P, N, F, kd=(0.4, 0.3, 0.2), quads=False, negative=False):
    """Faces as triangles (or pairs merged into quads when they share an edge
    in order, for the fan-triangulation test); indices 1-based or negative."""
    mtl = path.with_suffix(".mtl")
    mtl.write_text(f"newmtl skin\nKd {kd[0]} {kd[1]} {kd[2]}\nNs 250\nnewmtl dark\nKd 0.05 0.05 0.05\n")
    lines = [f"mtllib {mtl.name}"]
    lines += [f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}" for p in P]
    lines += [f"vn {n[0]:.9g} {n[1]:.9g} {n[2]:.9g}" for n in N]
    lines += [f"vt {0.5:.3f} {0.5:.3f}" for _ in P]
    nv = len(P)

    def ref(i):
        j = i - nv if negative else i + 1
        return f"{j}/{j}/{j}"
    lines.append("usemtl skin")
    half = len(F) // 2
    for k, tri in enumerate(F):
        if k == half:
            lines.append("usemtl dark")
        lines.append("f " + " ".join(ref(i) for i in tri))
    if quads:
        a, b, c = F[0]
        lines.append("usemtl skin")
        lines.append("f " + " ".join(ref(i) for i in (a, b, c, a)))  # degenerate quad -> 2 triangles
    path.write_text("\n".join(lines) + "\n")
