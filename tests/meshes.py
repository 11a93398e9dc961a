"""Small meshes used by the reference's own voxelize tests, rebuilt from their
definitions (tests/test_voxelize.cpp:17-88), plus exact-arithmetic helpers."""
from fractions import Fraction

import numpy as np


def subdivided_cube(n: int):
    """Unit cube with each face split into an n x n grid of quads, two triangles each
    (test_voxelize.cpp:20-52): watertight, convex, volume exactly 1."""
    vid = {}
    pts = []

    def vert(x, y, z):
        k = (x, y, z)
        if k not in vid:
            vid[k] = len(pts)
            pts.append((x / n, y / n, z / n))
        return vid[k]

    tris = []

    def quad(a, b, c, d):
        tris.append((a, b, c))
        tris.append((a, c, d))

    for i in range(n):
        for j in range(n):
            quad(vert(i, j, 0), vert(i + 1, j, 0), vert(i + 1, j + 1, 0), vert(i, j + 1, 0))
            quad(vert(i, j, n), vert(i, j + 1, n), vert(i + 1, j + 1, n), vert(i + 1, j, n))
            quad(vert(i, 0, j), vert(i, 0, j + 1), vert(i + 1, 0, j + 1), vert(i + 1, 0, j))
            quad(vert(i, n, j), vert(i + 1, n, j), vert(i + 1, n, j + 1), vert(i, n, j + 1))
            quad(vert(0, i, j), vert(0, i + 1, j), vert(0, i + 1, j + 1), vert(0, i, j + 1))
            quad(vert(n, i, j), vert(n, i, j + 1), vert(n, i + 1, j + 1), vert(n, i + 1, j))
    return np.array(pts, dtype=np.float64), np.array(tris, dtype=np.int32)


def uv_sphere(r: float, nlat: int, nlon: int):
    """Triangulated UV sphere (test_voxelize.cpp:56-88)."""
    pts = [(0.0, 0.0, r)]
    ids = {}
    for i in range(1, nlat):
        th = np.pi * i / nlat
        for j in range(nlon):
            ph = 2 * np.pi * j / nlon
            ids[(i, j)] = len(pts)
            pts.append((r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph),
                        r * np.cos(th)))
    bot = len(pts)
    pts.append((0.0, 0.0, -r))
    tris = []
    for j in range(nlon):
        tris.append((0, ids[(1, j)], ids[(1, (j + 1) % nlon)]))
    for i in range(1, nlat - 1):
        for j in range(nlon):
            j1 = (j + 1) % nlon
            tris.append((ids[(i, j)], ids[(i + 1, j)], ids[(i + 1, j1)]))
            tris.append((ids[(i, j)], ids[(i + 1, j1)], ids[(i, j1)]))
    for j in range(nlon):
        tris.append((ids[(nlat - 1, j)], bot, ids[(nlat - 1, (j + 1) % nlon)]))
    return np.array(pts, dtype=np.float64), np.array(tris, dtype=np.int32)


def brick(X, Y, Z):
    """Axis-aligned box [0,X]x[0,Y]x[0,Z] (bindings/python/tests/
    test_cybermesh_voxelize.py:32-41)."""
    pts = np.array([[0, 0, 0], [X, 0, 0], [X, Y, 0], [0, Y, 0],
                    [0, 0, Z], [X, 0, Z], [X, Y, Z], [0, Y, Z]], dtype=np.float64)
    tris = []
    for a, b, c, d in [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4),
                       (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]:
        tris.append((a, b, c))
        tris.append((a, c, d))
    return pts, np.array(tris, dtype=np.int32)


# ---- exact rational arithmetic -------------------------------------------------------

def det3(m):
    return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
            - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
            + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


def det4(m):
    s = 0
    for j in range(4):
        if m[0][j] == 0:
            continue
        sub = [[m[r][c] for c in range(4) if c != j] for r in range(1, 4)]
        s += (-1) ** j * m[0][j] * det3(sub)
    return s


def frac_orient3d(a, b, c, d) -> int:
    """Sign of det[a-d; b-d; c-d] in exact rational arithmetic (Shewchuk's orient3d:
    positive when d lies below the plane through a, b, c oriented counterclockwise
    seen from above)."""
    A = [Fraction(float(x)) for x in a]
    B = [Fraction(float(x)) for x in b]
    C = [Fraction(float(x)) for x in c]
    D = [Fraction(float(x)) for x in d]
    m = [[A[i] - D[i] for i in range(3)], [B[i] - D[i] for i in range(3)],
         [C[i] - D[i] for i in range(3)]]
    v = det3(m)
    return (v > 0) - (v < 0)


def frac_insphere(a, b, c, d, e) -> int:
    """Sign of the lifted 4x4 determinant [p-e, |p-e|^2] over p = a, b, c, d."""
    E = [Fraction(float(x)) for x in e]
    rows = []
    for p in (a, b, c, d):
        r = [Fraction(float(x)) - E[i] for i, x in enumerate(p)]
        rows.append(r + [r[0] * r[0] + r[1] * r[1] + r[2] * r[2]])
    v = det4(rows)
    return (v > 0) - (v < 0)


def box_on_cell_centres():
    """Box [1/8, 7/8] x [1/8, 5/8] x [1/8, 3/8] plus anchor points (0,0,0) and (1,1,1):
    on a 4^3 grid over [0, 1]^3 the cell centres (1/8 + k/4) lie exactly on its faces,
    edges and corners, where the distance is 0."""
    pts, tris = brick(0.75, 0.5, 0.25)
    pts = pts + 0.125
    pts = np.vstack([pts, [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]])
    return pts, tris

