"""Closed test meshes shared by the voxelizer tests: rotated boxes, random tetrahedra and an
icosphere (convex, outward winding) and an extruded L-shaped room."""
import numpy as np

from parallelfdtd_amd import geometry as G


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def icosphere(subdiv=1):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                cache[k] = len(v) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v), np.array(f, np.uint32)


def convex_shapes():
    rng = np.random.default_rng(42)
    out = []
    for k in range(3):  # rotated boxes
        b = G.box_mesh((-0.5, -0.35, -0.25), (0.5, 0.35, 0.25))
        v = b.vertices.astype(float) @ rotation(rng).T + 2.0
        out.append((f"box{k}", v, b.indices))
    for k in range(3):  # random tetrahedra, outward winding
        p = rng.uniform(0, 1, (4, 3))
        f = np.array([(0, 1, 2), (0, 3, 1), (0, 2, 3), (1, 3, 2)], np.uint32)
        if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0]) > 0:
            f = f[:, ::-1].copy()
        out.append((f"tet{k}", p, f))
    v, f = icosphere(2)
    out.append(("icosphere", v * 0.6 + 1.0, f))
    return out


def l_room(h=1.0):
    poly = [(0, 0), (2, 0), (2, 1), (1, 1), (1, 2), (0, 2)]
    n = len(poly)
    v = [(x, y, 0.0) for x, y in poly] + [(x, y, h) for x, y in poly]
    f = []
    for i in range(1, n - 1):  # fans from vertex 0 (the L is star-shaped about it)
        f.append((0, i + 1, i))
        f.append((n, n + i, n + i + 1))
    for i in range(n):
        j = (i + 1) % n
        f += [(i, j, n + j), (i, n + j, n + i)]
    return np.array(v, np.float32), np.array(f, np.uint32)
