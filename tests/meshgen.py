"""Independent numpy mesh generators for the tests.

They define the synthetic meshes of the workloads (icosphere, ellipsoid,
structured plate) once more, independently of the product's C++ generators,
so that tests can check the product's meshes element by element.

Conventions: triangles counter-clockwise seen from outside (outward normal
(B-A)x(C-B)), 0-based vertex ids.
"""
import numpy as np

_PHI = (1.0 + 5.0 ** 0.5) / 2.0

ICO_V = np.array([
    [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
    [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
    [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1],
], dtype=np.float64)

ICO_F = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
], dtype=np.int64)


def icosphere(subdiv, radius=1.0):
    """Icosahedron, 4-way subdivided `subdiv` times, vertices projected to the sphere.

    T = 20*4^s, V = 10*4^s+2.  Children of face (a,b,c) in order:
    (a,ab,ca), (b,bc,ab), (c,ca,bc), (ab,bc,ca); a new midpoint vertex is
    appended the first time its edge is met."""
    v = [tuple(p / np.linalg.norm(p)) for p in ICO_V]
    f = [tuple(t) for t in ICO_F]
    for _ in range(subdiv):
        mid = {}
        nf = []

        def m(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in mid:
                p = (np.array(v[a]) + np.array(v[b])) * 0.5
                p = p / np.linalg.norm(p)
                mid[key] = len(v)
                v.append(tuple(p))
            return mid[key]

        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    xyz = np.array(v, dtype=np.float64) * radius
    return xyz, np.array(f, dtype=np.int32)


def ellipsoid(subdiv, a, b, c):
    xyz, tri = icosphere(subdiv, 1.0)
    return xyz * np.array([a, b, c]), tri


def plate(n, size=1.0):
    """Square plate in z=0 of size x size, n x n cells, 2 triangles per cell (CCW seen from +z)."""
    h = size / n
    xs = np.arange(n + 1) * h - 0.5 * size
    xyz = np.array([[x, y, 0.0] for y in xs for x in xs])
    tri = []
    for j in range(n):
        for i in range(n):
            v00 = j * (n + 1) + i
            v10 = v00 + 1
            v01 = v00 + (n + 1)
            v11 = v01 + 1
            tri.append((v00, v10, v11))
            tri.append((v00, v11, v01))
    return xyz, np.array(tri, dtype=np.int32)


def rect(nx, ny, origin, ex, ey):
    """Parallelogram origin + s ex + t ey (s, t in [0, 1]) with nx x ny cells,
    two triangles per cell, counter-clockwise about ex x ey."""
    origin, ex, ey = (np.asarray(v, dtype=np.float64) for v in (origin, ex, ey))
    xyz = np.array([origin + (i / nx) * ex + (j / ny) * ey for j in range(ny + 1) for i in range(nx + 1)])
    tri = []
    for j in range(ny):
        for i in range(nx):
            v00 = j * (nx + 1) + i
            v10, v01 = v00 + 1, v00 + nx + 1
            tri.append((v00, v10, v01 + 1))
            tri.append((v00, v01 + 1, v01))
    return np.round(xyz, 9), np.array(tri, dtype=np.int32)


def join(parts):
    """Concatenate meshes; every part keeps its own vertices (no merging)."""
    xyz, tri, off = [], [], 0
    for x, t in parts:
        xyz.append(x)
        tri.append(t + off)
        off += len(x)
    return np.concatenate(xyz), np.concatenate(tri).astype(np.int32)


def cube_faces(res, size=1.0):
    """Closed cube surface of edge `size` centred at the origin whose six faces
    (+x, -x, +y, -y, +z, -z) are meshed separately with res[f] x res[f]
    cells and their own vertices, triangles counter-clockwise seen from
    outside.  Faces of equal resolution meet with coincident vertices under
    distinct ids, faces of different resolution nonconformally."""
    h = size / 2
    s = 2 * h
    faces = [((h, -h, -h), (0, s, 0), (0, 0, s)), ((-h, -h, -h), (0, 0, s), (0, s, 0)),
             ((-h, h, -h), (0, 0, s), (s, 0, 0)), ((-h, -h, -h), (s, 0, 0), (0, 0, s)),
             ((-h, -h, h), (s, 0, 0), (0, s, 0)), ((-h, -h, -h), (0, s, 0), (s, 0, 0))]
    return join([rect(r, r, o, a, b) for r, (o, a, b) in zip(res, faces)])


def plate_pair(nl, nr, size=1.0):
    """size x size plate in z = 0 cut along x = 0 into two halves meshed
    separately: the left half with nl cells along y (and nl // 2 + 1 along
    x), the right half with nr; the halves share no vertex ids."""
    h = size / 2
    left = rect(max(1, nl // 2 + 1), nl, (-h, -h, 0), (h, 0, 0), (0, size, 0))
    right = rect(max(1, nr // 2 + 1), nr, (0, -h, 0), (h, 0, 0), (0, size, 0))
    return join([left, right])
