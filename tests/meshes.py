"""Small test meshes built independently of the engine's generators."""
import itertools

import numpy as np

# Kuhn split of the unit cube into 6 tets, one per monotone corner path from
# corner 0 to corner 7 (corner b has x = bit 2, y = bit 1, z = bit 0); every
# cube uses the same split, so the mesh is conforming.
_KUHN = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]


def cube_grid(n, lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0), field=lambda x: x[:, 2]):
    """n^3 cubes, 6 tets each, pressure given by `field` at the vertices."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    V = np.array([lo + (hi - lo) * np.array([i, j, k]) / n for i in range(n + 1) for j in range(n + 1)
                  for k in range(n + 1)])
    T = []
    for i, j, k in itertools.product(range(n), repeat=3):
        corner = [idx(i + (b >> 2 & 1), j + (b >> 1 & 1), k + (b & 1)) for b in range(8)]
        for t in _KUHN:
            q = [corner[v] for v in t]
            a, b, c, d = V[q]
            if np.dot(b - a, np.cross(c - a, d - a)) < 0:
                q[2], q[3] = q[3], q[2]
            T.append(q)
    return V, np.array(T, dtype=np.int32), field(V).astype(np.float64)


def signed_volumes(V, T):
    a, b, c, d = (V[T[:, i]] for i in range(4))
    return np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)) / 6.0


def pose(R=np.eye(3), t=(0, 0, 0)):
    X = np.zeros(12)
    X[:9] = np.asarray(R, float).reshape(9)
    X[9:] = t
    return X


def rot(axis, ang):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
