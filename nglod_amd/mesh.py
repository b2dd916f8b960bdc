"""Training-target mesh: level-4 icosphere with a seeded radial displacement,
normalized as sdf-net/lib/torchgp/normalize.py:24-38 does (centre on the
bounding-box centre, scale to unit maximum radius).

The reference loads an .obj with tinyobjloader (sdf-net/lib/torchgp/load_obj.py);
the benchmark mesh is generated in-process instead, deterministically.
"""
import numpy as np
import torch


def icosphere(level=4):
    """Vertices [V, 3] (float64, unit sphere) and faces [F, 3] (int64) of a
    subdivided icosahedron: level 4 gives 2562 vertices and 5120 faces."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    F = list(faces)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    return np.stack(V), np.array(F, dtype=np.int64)


def radial_noise(V, seed=0, amplitude=0.1, waves=8):
    """Smooth seeded noise on the unit sphere in [-amplitude, amplitude]: a sum
    of `waves` cosine waves along random directions with random frequencies and
    phases, normalized by its maximum over the vertices."""
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(waves, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    freq = rng.uniform(1.0, 4.0, size=waves)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=waves)
    amp = rng.uniform(0.5, 1.0, size=waves)
    n = (amp[None] * np.cos(V @ dirs.T * freq[None] + phase[None])).sum(1)
    return amplitude * n / np.abs(n).max()


def normalize(V):
    """Centre on the bounding-box centre and scale to unit maximum radius
    (sdf-net/lib/torchgp/normalize.py:24-38)."""
    V = V - (V.max(0) + V.min(0)) / 2.0
    return V / np.sqrt((V ** 2).sum(1).max())


def displaced_icosphere(seed=0, level=4, amplitude=0.1):
    """(V [2562, 3] float32, F [5120, 3] int64) of the benchmark shape."""
    V, F = icosphere(level)
    V = V * (1.0 + radial_noise(V, seed, amplitude))[:, None]
    return normalize(V).astype(np.float32), F


def triangles(V, F):
    """Triangle soup [F, 3, 3] (the `mesh = V[F]` of MeshDataset.py:72)."""
    return V[F]


def face_cdf(tris):
    """Inclusive cumulative area distribution over faces, float64, last = 1
    (area_weighted_distribution.py:36-45: areas from |cross| / 2, normalized)."""
    t = tris.astype(np.float64)
    a = t[:, 0] - t[:, 1]
    b = t[:, 1] - t[:, 2]
    area = np.linalg.norm(np.cross(a, b), axis=1) * 0.5
    c = np.cumsum(area / (area.sum() + 1e-10))
    c[-1] = max(c[-1], 1.0)
    return c
